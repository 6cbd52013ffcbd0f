// implicit.hip -- the dual-time implicit update on gfx950 (FP64).
//
// Dual_time_stepping.cpp:18-1794 split into device passes:
//   k_zero_ghosts     NL/NR ghost planes (198-323)
//   k_sponge_*        absorbing-zone sources (503-664), optional
//   k_rhs_diag        BDF2 + flux divergence + viscous + body force, Gamma^-1,
//                     spectral-radius diagonal, NL = NR = D MR (737-908)
//   k_lusgs_plane     one i+j+k hyperplane of the forward / backward sweep
//                     (983-1300, 1312-1664)
//   k_update_norm     primitive update and per-block partial norms (1702-1782)
//   k_finish_norm     fixed-order sum of the partials, sqrt, scale
// plus the Jacobi sweep of DPLUSGS.cpp:544-1125.
#include <vector>

#include "cgli_internal.hpp"

namespace cgli {

struct CP { double rho, u, v, w, VV, P, C, T, H, beta; };

// Dual_time_stepping.cpp:759-769, from the cell's conserved variables and
// J.  Quotients that share a divisor (U1, rho) use its reciprocal and
// div_r: the division's bits, four divisions instead of seven.
__device__ __forceinline__ CP cell_prim_v(const double U[5], double Jc, double K, double e)
{
    CP q;
    const double U1 = U[0], iU1 = 1.0 / U1;
    q.rho = U1 * Jc;
    q.u = div_r(U[1], U1, iU1);
    q.v = div_r(U[2], U1, iU1);
    q.w = div_r(U[3], U1, iU1);
    q.VV = q.u * q.u + q.v * q.v + q.w * q.w;
    q.P = (U[4] * Jc - 0.5 * q.rho * q.VV) * (K - 1);
    const double irho = 1.0 / q.rho;
    q.C = div_r(K * q.P, q.rho, irho);
    q.T = div_r(q.P, q.rho, irho);
    q.H = 0.5 * q.VV + q.C / (K - 1);
    const double r = q.VV / q.C;
    q.beta = (r > e) ? r : e;
    return q;
}

__device__ __forceinline__ CP cell_prim(const C5 &U, const double *__restrict__ J, long c, double K, double e)
{
    const double u[5] = {U.v[0][c], U.v[1][c], U.v[2][c], U.v[3][c], U.v[4][c]};
    return cell_prim_v(u, J[c], K, e);
}

// split coefficients of one neighbour along one direction
// (Dual_time_stepping.cpp:1054-1170); late52: B/C ordering of x52..x54
struct SC { double p11, p12, p13, p14, p21, p22, p31, p41, p51, p52, p53, p54, m11, m22; };

__device__ __forceinline__ SC split(const CP &q, double mx, double my, double mz, double K, bool late52)
{
    SC s;
    const double b = q.beta;
    const double Un = mx * q.u + my * q.v + mz * q.w;
    const double S = sqrt(Un * Un * (b - 1) * (b - 1) + 4 * b * q.C * (mx * mx + my * my + mz * mz));
    const double lam = ((b + 1) * fabs(Un) + S) / 2;
    const double d = fabs(lam);
    s.p11 = 0.5 * (b * Un + d);
    s.p12 = 0.5 * q.P * K * b * mx;
    s.p13 = 0.5 * q.P * K * b * my;
    s.p14 = 0.5 * q.P * K * b * mz;
    s.p21 = 0.5 * mx / q.rho;
    s.p22 = 0.5 * (Un + d);   // also x33, x44, x55
    s.p31 = 0.5 * my / q.rho;
    s.p41 = 0.5 * mz / q.rho;
    s.p51 = 0.5 * Un * (K - 1) * (b - 1) / q.rho / K;
    if (!late52) {
        s.p52 = 0.5 * mx * b * q.T * (K - 1);
        s.p53 = 0.5 * my * b * q.T * (K - 1);
        s.p54 = 0.5 * mz * b * q.T * (K - 1);
    } else {
        s.p52 = 0.5 * b * q.T * (K - 1) * mx;
        s.p53 = 0.5 * b * q.T * (K - 1) * my;
        s.p54 = 0.5 * b * q.T * (K - 1) * mz;
    }
    s.m11 = 0.5 * (b * Un - d);
    s.m22 = 0.5 * (Un - d);
    return s;
}

struct DG { double d11, d22, d51; };

// Dual_time_stepping.cpp:856-895
__device__ __forceinline__ DG diag(const CP &q, const double m[9], const Phys &p, double deltaT)
{
    const double K = p.K, b = q.beta;
    const double U = m[0] * q.u + m[1] * q.v + m[2] * q.w;
    const double V = m[3] * q.u + m[4] * q.v + m[5] * q.w;
    const double W = m[6] * q.u + m[7] * q.v + m[8] * q.w;
    const double Sx = sqrt(U * U * (b - 1) * (b - 1) + 4 * b * q.C * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]));
    const double Sy = sqrt(V * V * (b - 1) * (b - 1) + 4 * b * q.C * (m[3] * m[3] + m[4] * m[4] + m[5] * m[5]));
    const double Sz = sqrt(W * W * (b - 1) * (b - 1) + 4 * b * q.C * (m[6] * m[6] + m[7] * m[7] + m[8] * m[8]));
    const double lX = ((b + 1) * fabs(U) + Sx) / 2;
    const double lY = ((b + 1) * fabs(V) + Sy) / 2;
    const double lZ = ((b + 1) * fabs(W) + Sz) / 2;
    const double lambda = lX / p.dxi + lY / p.deta + lZ / p.dzeta;
    DG D;
    D.d11 = 2 * deltaT / (3 * b + 2 * lambda * deltaT);
    D.d22 = 2 * deltaT / (3 + 2 * lambda * deltaT);
    const double t1 = 6 * (K - 1) * (b - 1) * deltaT;
    const double t2 = K * (3 + 2 * lambda * deltaT) * (3 * b + 2 * lambda * deltaT) * q.rho;
    D.d51 = -t1 / t2;
    return D;
}

__device__ __forceinline__ void load9(const M9 &M, long c, double m[9])
{
#pragma unroll
    for (int q = 0; q < 9; q++) m[q] = M.m[q][c];
}

// ------------------------------------------------------------------ ghosts
__global__ void k_zero_ghosts(Geo g, F5 NL, F5 NR, int *__restrict__ counters)
{
    // every (i,j,k) of i=2..gend+1, j=1..NYY, k=1..NZZ that is a ghost of
    // the increment arrays, one shell per blockIdx.y: 0 the planes j=1 and
    // j=NYY, 1 the planes k=1 and k=NZZ, 2 the x planes i=2 (first slab or
    // periodic) and i=gend+1 (last slab or periodic).  counters (the
    // wavefront's ticket and error flag) are zeroed on the way.
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (counters && t < 2 && blockIdx.y == 0) counters[t] = 0;
    const int ni = g.gend, nj = g.NYY, nk = g.NZZ;
    int i, j, k;
    if (blockIdx.y == 0) {
        if (t >= 2L * ni * nk) return;
        k = 1 + (int)(t % nk);
        const long r = t / nk;
        i = 2 + (int)(r % ni);
        j = (r / ni) ? g.NYY : 1;
    } else if (blockIdx.y == 1) {
        if (t >= 2L * ni * nj) return;
        j = 1 + (int)(t % nj);
        const long r = t / nj;
        i = 2 + (int)(r % ni);
        k = (r / ni) ? g.NZZ : 1;
    } else {
        if (t >= 2L * nj * nk) return;
        k = 1 + (int)(t % nk);
        const long r = t / nk;
        j = 1 + (int)(r % nj);
        const bool hi = (r / nj) != 0;
        if (hi ? !(g.last || g.xper) : !(g.first || g.xper)) return;
        i = hi ? g.gend + 1 : 2;
    }
    const long c = ix(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        NL.v[v][c] = 0;
        NR.v[v][c] = 0;
    }
}

void launch_zero_increment_ghosts(const Geo &g, F5 NL, F5 NR, int *counters, hipStream_t s)
{
    const long n = std::max(2L * g.gend * std::max(g.NZZ, g.NYY), 2L * g.NYY * g.NZZ);
    hipLaunchKernelGGL(k_zero_ghosts, dim3((unsigned)((n + 255) / 256), 3), dim3(256), 0, s, g, NL, NR, counters);
}

__global__ void k_zero(double *p, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

void launch_zero(double *p, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

// ------------------------------------------------------------ reductions
// Block sums in a fixed order: wave shuffles then a serial pass over the
// waves, so two runs give identical bits.
template <int NV>
__device__ __forceinline__ void block_sum(double (&x)[NV], double *out)
{
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double a = x[v];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0) red[v][wid] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += red[v][w];
            out[v] = s;
        }
    }
}

// ----------------------------------------------------------------- sponge
// targetU integral over the inlet plane i=3 (Dual_time_stepping.cpp:503-514)
__global__ void __launch_bounds__(256) k_sponge_target(Geo g, Phys p, C5 U, const double *__restrict__ etdy,
                                                       double *partials)
{
    const int nk = g.NZZ - 1, nj = g.NYY - 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double x[1] = {0.0};
    if (t < (long)nk * nj) {
        const int j = 2 + (int)(t / nk), k = 2 + (int)(t % nk);
        const long c = ix(g, 3, j, k);
        const double u = U.v[1][c] / U.v[0][c];
        x[0] = u * p.deta / etdy[c] * p.dzeta;
    }
    double o[1];
    block_sum<1>(x, o);
    if (threadIdx.x == 0) partials[blockIdx.x] = o[0];
}

__global__ void k_sponge_finish(int nb, const double *partials, double DZ, double *targetU)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double z = 0.0;
        for (int b = 0; b < nb; b++) z += partials[b];
        *targetU = z * DZ;
    }
}

struct SpongeArgs {
    int gstart, gend0, nx_global, nx_inlet, nx_outlet;
    int in_lo, in_hi, out_lo, out_hi;   // i ranges (empty when lo > hi)
};

__global__ void __launch_bounds__(256) k_sponge_res(Geo g, Phys p, double e, SpongeArgs a, C5 U,
                                                   const double *__restrict__ J, const double *targetU_p, F5 Res)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
    const bool out = (i >= a.out_lo && i <= a.out_hi);
    const bool in = (i >= a.in_lo && i <= a.in_hi);
    if (!out && !in) {
#pragma unroll
        for (int v = 0; v < 5; v++) Res.v[v][c] = 0;
        return;
    }
    const double K = p.K;
    const long im = c - g.Ym * g.Zm;
    const double Jc = J[c];
    const double rho = U.v[0][c] * Jc;
    const double u = U.v[1][c] / U.v[0][c];
    const double v = U.v[2][c] / U.v[0][c];
    const double w = U.v[3][c] / U.v[0][c];
    const double VV = u * u + v * v + w * w;
    const double P = (U.v[4][c] * Jc - 0.5 * rho * VV) * (K - 1);
    const double C = K * P / rho;
    const double r = VV / C;
    const double beta = (r > e) ? r : e;
    const double C_plan = 0.5 * sqrt(u * u * (beta - 1) * (beta - 1) + 4 * beta * C);
    double s1, sg, t[5];
    if (out) {   // the outlet loop runs after the inlet loop and overwrites it
        const double targetU = *targetU_p;
        const int nxr = a.nx_global + 1;
        s1 = pow((((i + a.gstart - 0.5 - (nxr - a.nx_outlet)) * 1.0) / a.nx_outlet), 3.0) * 1.0 * C_plan;
        sg = 1.25 * s1 / p.high;
        t[0] = rho - p.rho_ref;
        t[1] = rho * u - p.rho_ref * targetU;
        t[2] = rho * v;
        t[3] = rho * w;
        t[4] = U.v[4][c] * Jc - p.E_ref - 0.5 * p.rho_ref * targetU * targetU;
    } else {
        s1 = pow((((a.nx_inlet - a.gstart - 0.5) * 1.0) / a.nx_inlet), 3.0) * 1.0 * C_plan;
        sg = 0.035 * s1 / p.high;
        t[0] = rho - p.rho_ref;
        t[1] = rho * u;
        t[2] = rho * v;
        t[3] = rho * w;
        t[4] = U.v[4][c] * Jc - p.E_ref;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) Res.v[q][c] = -s1 * (U.v[q][c] * Jc - U.v[q][im] * J[im]) / p.dxi - sg * t[q];
}

void launch_sponge(const Geo &g, const Phys &p, double e, int gstart, int gend0, int nx_global, int nx_inlet,
                   int nx_outlet, C5 U, const double *J, const double *etdy, F5 Res, double *scratch,
                   hipStream_t s)
{
    const long n = (long)(g.NZZ - 1) * (g.NYY - 1);
    const int nb = (int)((n + 255) / 256);
    double *targetU = scratch;
    double *partials = scratch + 1;
    if (g.first) {
        hipLaunchKernelGGL(k_sponge_target, dim3(nb), dim3(256), 0, s, g, p, U, etdy, partials);
        const double DZ = 1. / (p.high * p.dzeta * g.nz);
        hipLaunchKernelGGL(k_sponge_finish, dim3(1), dim3(64), 0, s, nb, partials, DZ, targetU);
    } else {
        launch_zero(targetU, 1, s);
    }
    SpongeArgs a{gstart, gend0, nx_global, nx_inlet, nx_outlet, 1, 0, 1, 0};
    if (gstart < nx_inlet) {
        a.in_lo = 3;
        a.in_hi = (gend0 < nx_inlet) ? g.gend : nx_inlet - gstart + 2;
    }
    if (gend0 > (nx_global - nx_outlet - 1)) {
        a.out_lo = (gstart > (nx_global - nx_outlet - 1)) ? 3 : (nx_global - nx_outlet) - gstart + 2;
        a.out_hi = g.gend;
    }
    dim3 b(64, 4);
    dim3 gr((g.NZ - 1 + b.x - 1) / b.x, (g.NY - 1 + b.y - 1) / b.y, g.nx);
    hipLaunchKernelGGL(k_sponge_res, gr, b, 0, s, g, p, e, a, U, J, targetU, Res);
}

// --------------------------------------------------------------- RHS+diag
// Reciprocals of the constant divisors, formed on the host (correctly
// rounded) for div_r.
struct RInv { double dt2, dxi, deta, dzeta; };

// INC: also the starting increments NL = NR = D MR (the schedules whose
// sweeps read them from the fields; schedule 4 keeps its own copies, and
// then neither the diagonal nor the metrics are touched here).  WR: the
// residual source Res / J is added.  Every load of the cell is issued before
// the first is used (the kernel is one round trip to memory), and each
// division of Dual_time_stepping.cpp:773-828 is div_r with its divisor's
// reciprocal: MR has the bits of the plain divisions.
template <bool INC, bool WR>
__global__ void __launch_bounds__(256) k_rhs_diag(Geo g, Phys p, double deltaT, RInv ri, double e, C5 U_,
                                                  C5 U, C5 Uq, const double *__restrict__ J, M9 met, C5 Fx,
                                                  C5 Fy, C5 Fz, C4 vF, C5 Res, F5 MR, F5 NL, F5 NR)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const double K = p.K;
    const long c = ix(g, i, j, k);
    const long sx = g.Ym * g.Zm, sy = g.Zm;
    // flux face indices: x faces of cell (i,j,k) at [i][j-1][k-1] and
    // [i-1][j-1][k-1]; y at [i-1][j][k-1] and [i-1][j-1][k-1]; z at
    // [i-1][j-1][k] and [i-1][j-1][k-1] (Dual_time_stepping.cpp:810-828)
    const long f0 = c - sx - sy - 1;
    double u_[5], un[5], uq[5], fx[5][2], fy[5][2], fz[5][2], vf[4], rs[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
        u_[n] = U_.v[n][c];
        un[n] = U.v[n][c];
        uq[n] = Uq.v[n][c];
        fx[n][0] = Fx.v[n][f0];
        fx[n][1] = Fx.v[n][f0 + sx];
        fy[n][0] = Fy.v[n][f0];
        fy[n][1] = Fy.v[n][f0 + sy];
        fz[n][0] = Fz.v[n][f0];
        fz[n][1] = Fz.v[n][f0 + 1];
        rs[n] = WR ? Res.v[n][c] : 0.0;
    }
#pragma unroll
    for (int n = 0; n < 4; n++) vf[n] = vF.v[n][c];
    const double Jc = J[c];
    __builtin_amdgcn_sched_barrier(0);   // the loads above stay ahead of all arithmetic
    const CP q = cell_prim_v(u_, Jc, K, e);
    const double rho = q.rho, u = q.u, v = q.v, w = q.w, VV = q.VV, T = q.T, H = q.H, beta = q.beta;
    const double temp = rho * K, irho = 1.0 / rho, itemp = 1.0 / temp, iJc = 1.0 / Jc;
    double d[5][5];
    d[0][0] = -(H * (K - 1) + VV - K * (T + VV)) * beta;
    d[0][1] = -(K - 1) * u * beta;
    d[0][2] = -(K - 1) * v * beta;
    d[0][3] = -(K - 1) * w * beta;
    d[0][4] = (K - 1) * beta;
    d[1][0] = div_r(-u, rho, irho); d[1][1] = irho; d[1][2] = 0; d[1][3] = 0; d[1][4] = 0;
    d[2][0] = div_r(-v, rho, irho); d[2][1] = 0; d[2][2] = irho; d[2][3] = 0; d[2][4] = 0;
    d[3][0] = div_r(-w, rho, irho); d[3][1] = 0; d[3][2] = 0; d[3][3] = irho; d[3][4] = 0;
    d[4][0] = div_r((K - 1) * (VV + K * T * beta + (K - 1) * VV * beta + H * (-1 + beta - K * beta)), temp, itemp);
    d[4][1] = div_r(-(K - 1) * u * (1 + (K - 1) * beta), temp, itemp);
    d[4][2] = div_r(-(K - 1) * v * (1 + (K - 1) * beta), temp, itemp);
    d[4][3] = div_r(-(K - 1) * w * (1 + (K - 1) * beta), temp, itemp);
    d[4][4] = div_r((K - 1) * (1 + (K - 1) * beta), temp, itemp);
    const double f = -(rho - p.rho_ref) * p.grav + p.f_const;
    const double dt2 = 2 * deltaT;
    double Rk[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
        const double Rp = div_r(-(3 * u_[n] - 4 * un[n] + uq[n]), dt2, ri.dt2);
        const double Rf = -(div_r(fx[n][1] - fx[n][0], p.dxi, ri.dxi) + div_r(fy[n][1] - fy[n][0], p.deta, ri.deta) +
                            div_r(fz[n][1] - fz[n][0], p.dzeta, ri.dzeta));
        double r = WR ? div_r(rs[n], Jc, iJc) + Rp + Rf : Rp + Rf;
        if (n >= 1) r = r + vf[n - 1];
        if (n == 1) r = r + div_r(f, Jc, iJc);
        if (n == 4) r = r + div_r(f * u, Jc, iJc);
        Rk[n] = r;
    }
    double mr[5];
#pragma unroll
    for (int a = 0; a < 5; a++)
        mr[a] = (d[a][0] * Rk[0] + d[a][1] * Rk[1] + d[a][2] * Rk[2] + d[a][3] * Rk[3] + d[a][4] * Rk[4]);
#pragma unroll
    for (int a = 0; a < 5; a++) MR.v[a][c] = mr[a];
    if constexpr (INC) {
        double m[9];
        load9(met, c, m);
        const DG D = diag(q, m, p, deltaT);
        double nl[5];
        nl[0] = D.d11 * mr[0];
        nl[1] = D.d22 * mr[1];
        nl[2] = D.d22 * mr[2];
        nl[3] = D.d22 * mr[3];
        nl[4] = D.d51 * mr[0] + D.d22 * mr[4];
#pragma unroll
        for (int a = 0; a < 5; a++) {
            NL.v[a][c] = nl[a];
            NR.v[a][c] = nl[a];
        }
    }
}

void launch_rhs_diag(const Geo &g, const Phys &p, double deltaT, double e, int with_res, C5 U_, C5 U, C5 Uq,
                     const double *J, M9 met, C5 Fx, C5 Fy, C5 Fz, C4 vF, C5 Res, F5 MR, F5 NL, F5 NR,
                     hipStream_t s, int write_inc)
{
    // a block spans the whole k row where it fits (255 cells at NZ = 256: one
    // block per row, every wave reading whole lines of the field layout)
    const int bx = g.NZ - 1 > 128 ? 256 : g.NZ - 1 > 64 ? 128 : 64;
    dim3 b(bx, 256 / bx);
    dim3 gr((g.NZ - 1 + b.x - 1) / b.x, (g.NY - 1 + b.y - 1) / b.y, g.nx);
    const RInv ri{1.0 / (2 * deltaT), 1.0 / p.dxi, 1.0 / p.deta, 1.0 / p.dzeta};
    auto kern = write_inc ? (with_res ? k_rhs_diag<true, true> : k_rhs_diag<true, false>)
                          : (with_res ? k_rhs_diag<false, true> : k_rhs_diag<false, false>);
    hipLaunchKernelGGL(kern, gr, b, 0, s, g, p, deltaT, ri, e, U_, U, Uq, J, met, Fx, Fy, Fz, vF, Res, MR, NL, NR);
}

// ------------------------------------------------------------- LU-SGS
__device__ __forceinline__ SC nb_split(const C5 &U_, const double *J, const M9 &met, long c, int dir,
                                       double K, double e)
{
    const CP q = cell_prim(U_, J, c, K, e);
    return split(q, met.m[3 * dir][c], met.m[3 * dir + 1][c], met.m[3 * dir + 2][c], K, dir != 0);
}

// forward: NL = D (MR - L NL), L from the lower neighbours (1172-1293)
__device__ __forceinline__ void fwd_cell(const Geo &g, const Phys &p, double deltaT, double e, long c,
                                         const C5 &U_, const double *J, const M9 &met, const C5 &MR, F5 &NL)
{
    const double K = p.K;
    const long ia = c - g.Ym * g.Zm, jb = c - g.Zm, kc = c - 1;
    const SC A = nb_split(U_, J, met, ia, 0, K, e);
    const SC B = nb_split(U_, J, met, jb, 1, K, e);
    const SC Cc = nb_split(U_, J, met, kc, 2, K, e);
    const double *X0 = NL.v[0], *X1 = NL.v[1], *X2 = NL.v[2], *X3 = NL.v[3], *X4 = NL.v[4];
    const double LL0 = (-A.p11 * X0[ia] - A.p12 * X1[ia] - A.p13 * X2[ia] - A.p14 * X3[ia]) / p.dxi +
                       (-B.p11 * X0[jb] - B.p12 * X1[jb] - B.p13 * X2[jb] - B.p14 * X3[jb]) / p.deta +
                       (-Cc.p11 * X0[kc] - Cc.p12 * X1[kc] - Cc.p13 * X2[kc] - Cc.p14 * X3[kc]) / p.dzeta;
    const double LL1 = (-A.p21 * X0[ia] - A.p22 * X1[ia]) / p.dxi + (-B.p21 * X0[jb] - B.p22 * X1[jb]) / p.deta +
                       (-Cc.p21 * X0[kc] - Cc.p22 * X1[kc]) / p.dzeta;
    const double LL2 = (-A.p31 * X0[ia] - A.p22 * X2[ia]) / p.dxi + (-B.p31 * X0[jb] - B.p22 * X2[jb]) / p.deta +
                       (-Cc.p31 * X0[kc] - Cc.p22 * X2[kc]) / p.dzeta;
    const double LL3 = (-A.p41 * X0[ia] - A.p22 * X3[ia]) / p.dxi + (-B.p41 * X0[jb] - B.p22 * X3[jb]) / p.deta +
                       (-Cc.p41 * X0[kc] - Cc.p22 * X3[kc]) / p.dzeta;
    const double LL4 =
        (-A.p51 * X0[ia] - A.p52 * X1[ia] - A.p53 * X2[ia] - A.p54 * X3[ia] - A.p22 * X4[ia]) / p.dxi +
        (-B.p51 * X0[jb] - B.p52 * X1[jb] - B.p53 * X2[jb] - B.p54 * X3[jb] - B.p22 * X4[jb]) / p.deta +
        (-Cc.p51 * X0[kc] - Cc.p52 * X1[kc] - Cc.p53 * X2[kc] - Cc.p54 * X3[kc] - Cc.p22 * X4[kc]) / p.dzeta;
    const CP q = cell_prim(U_, J, c, K, e);
    double m[9];
    load9(met, c, m);
    const DG D = diag(q, m, p, deltaT);
    NL.v[0][c] = D.d11 * (MR.v[0][c] - LL0);
    NL.v[1][c] = D.d22 * (MR.v[1][c] - LL1);
    NL.v[2][c] = D.d22 * (MR.v[2][c] - LL2);
    NL.v[3][c] = D.d22 * (MR.v[3][c] - LL3);
    NL.v[4][c] = D.d51 * (MR.v[0][c] - LL0) + D.d22 * (MR.v[4][c] - LL4);
}

// backward: NR = NL - D U NR, U from the upper neighbours (1355-1655)
__device__ __forceinline__ void bwd_cell(const Geo &g, const Phys &p, double deltaT, double e, long c,
                                         const C5 &U_, const double *J, const M9 &met, const F5 &NL, F5 &NR)
{
    const double K = p.K;
    const long ia = c + g.Ym * g.Zm, jb = c + g.Zm, kc = c + 1;
    const SC A = nb_split(U_, J, met, ia, 0, K, e);
    const SC B = nb_split(U_, J, met, jb, 1, K, e);
    const SC Cc = nb_split(U_, J, met, kc, 2, K, e);
    const double *X0 = NR.v[0], *X1 = NR.v[1], *X2 = NR.v[2], *X3 = NR.v[3], *X4 = NR.v[4];
    const double UU0 = (+A.m11 * X0[ia] + A.p12 * X1[ia] + A.p13 * X2[ia] + A.p14 * X3[ia]) / p.dxi +
                       (+B.m11 * X0[jb] + B.p12 * X1[jb] + B.p13 * X2[jb] + B.p14 * X3[jb]) / p.deta +
                       (+Cc.m11 * X0[kc] + Cc.p12 * X1[kc] + Cc.p13 * X2[kc] + Cc.p14 * X3[kc]) / p.dzeta;
    const double UU1 = (+A.p21 * X0[ia] + A.m22 * X1[ia]) / p.dxi + (+B.p21 * X0[jb] + B.m22 * X1[jb]) / p.deta +
                       (+Cc.p21 * X0[kc] + Cc.m22 * X1[kc]) / p.dzeta;
    const double UU2 = (+A.p31 * X0[ia] + A.m22 * X2[ia]) / p.dxi + (+B.p31 * X0[jb] + B.m22 * X2[jb]) / p.deta +
                       (+Cc.p31 * X0[kc] + Cc.m22 * X2[kc]) / p.dzeta;
    const double UU3 = (+A.p41 * X0[ia] + A.m22 * X3[ia]) / p.dxi + (+B.p41 * X0[jb] + B.m22 * X3[jb]) / p.deta +
                       (+Cc.p41 * X0[kc] + Cc.m22 * X3[kc]) / p.dzeta;
    const double UU4 =
        (+A.p51 * X0[ia] + A.p52 * X1[ia] + A.p53 * X2[ia] + A.p54 * X3[ia] + A.m22 * X4[ia]) / p.dxi +
        (+B.p51 * X0[jb] + B.p52 * X1[jb] + B.p53 * X2[jb] + B.p54 * X3[jb] + B.m22 * X4[jb]) / p.deta +
        (+Cc.p51 * X0[kc] + Cc.p52 * X1[kc] + Cc.p53 * X2[kc] + Cc.p54 * X3[kc] + Cc.m22 * X4[kc]) / p.dzeta;
    const CP q = cell_prim(U_, J, c, K, e);
    double m[9];
    load9(met, c, m);
    const DG D = diag(q, m, p, deltaT);
    NR.v[0][c] = NL.v[0][c] - D.d11 * UU0;
    NR.v[1][c] = NL.v[1][c] - D.d22 * UU1;
    NR.v[2][c] = NL.v[2][c] - D.d22 * UU2;
    NR.v[3][c] = NL.v[3][c] - D.d22 * UU3;
    NR.v[4][c] = NL.v[4][c] - (D.d51 * UU0 + D.d22 * UU4);
}

__global__ void __launch_bounds__(256) k_lusgs_plane(Geo g, Phys p, double deltaT, double e, int forward,
                                                     int first_cell, int count, const int *__restrict__ ijk,
                                                     C5 U_, const double *__restrict__ J, M9 met, C5 MR, F5 NL,
                                                     F5 NR)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int mi = first_cell + t;
    const int i = ijk[3 * mi], j = ijk[3 * mi + 1], k = ijk[3 * mi + 2];
    const long c = ix(g, i, j, k);
    if (forward) fwd_cell(g, p, deltaT, e, c, U_, J, met, MR, NL);
    else bwd_cell(g, p, deltaT, e, c, U_, J, met, NL, NR);
}

void launch_lusgs_plane(const Geo &g, const Phys &p, double deltaT, double e, int forward, int m, int first_cell,
                        int count, const int *ijk, C5 U_, const double *J, M9 met, C5 MR, F5 NL, F5 NR,
                        hipStream_t s)
{
    (void)m;
    const int b = 128;
    hipLaunchKernelGGL(k_lusgs_plane, dim3((count + b - 1) / b), dim3(b), 0, s, g, p, deltaT, e, forward,
                       first_cell, count, ijk, U_, J, met, MR, NL, NR);
}

// ------------------------------------------------------------ DP-LUR
// NRnew = D (MR - L NLold - U NLold) (DPLUSGS.cpp:575-1125), using the
// evident neighbour j+1 / k+1 for the upper part (the reference reads
// never-assigned j_ / k_ there)
__global__ void __launch_bounds__(256) k_dplur(Geo g, Phys p, double deltaT, double e, C5 U_,
                                               const double *__restrict__ J, M9 met, C5 MR, C5 NLo, F5 NRn)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const double K = p.K;
    const long c = ix(g, i, j, k);
    const long sx = g.Ym * g.Zm, sy = g.Zm;
    const double *X0 = NLo.v[0], *X1 = NLo.v[1], *X2 = NLo.v[2], *X3 = NLo.v[3], *X4 = NLo.v[4];
    double LL[5], UU[5];
    {
        const long ia = c - sx, jb = c - sy, kc = c - 1;
        const SC A = nb_split(U_, J, met, ia, 0, K, e);
        const SC B = nb_split(U_, J, met, jb, 1, K, e);
        const SC Cc = nb_split(U_, J, met, kc, 2, K, e);
        LL[0] = (-A.p11 * X0[ia] - A.p12 * X1[ia] - A.p13 * X2[ia] - A.p14 * X3[ia]) / p.dxi +
                (-B.p11 * X0[jb] - B.p12 * X1[jb] - B.p13 * X2[jb] - B.p14 * X3[jb]) / p.deta +
                (-Cc.p11 * X0[kc] - Cc.p12 * X1[kc] - Cc.p13 * X2[kc] - Cc.p14 * X3[kc]) / p.dzeta;
        LL[1] = (-A.p21 * X0[ia] - A.p22 * X1[ia]) / p.dxi + (-B.p21 * X0[jb] - B.p22 * X1[jb]) / p.deta +
                (-Cc.p21 * X0[kc] - Cc.p22 * X1[kc]) / p.dzeta;
        LL[2] = (-A.p31 * X0[ia] - A.p22 * X2[ia]) / p.dxi + (-B.p31 * X0[jb] - B.p22 * X2[jb]) / p.deta +
                (-Cc.p31 * X0[kc] - Cc.p22 * X2[kc]) / p.dzeta;
        LL[3] = (-A.p41 * X0[ia] - A.p22 * X3[ia]) / p.dxi + (-B.p41 * X0[jb] - B.p22 * X3[jb]) / p.deta +
                (-Cc.p41 * X0[kc] - Cc.p22 * X3[kc]) / p.dzeta;
        LL[4] = (-A.p51 * X0[ia] - A.p52 * X1[ia] - A.p53 * X2[ia] - A.p54 * X3[ia] - A.p22 * X4[ia]) / p.dxi +
                (-B.p51 * X0[jb] - B.p52 * X1[jb] - B.p53 * X2[jb] - B.p54 * X3[jb] - B.p22 * X4[jb]) / p.deta +
                (-Cc.p51 * X0[kc] - Cc.p52 * X1[kc] - Cc.p53 * X2[kc] - Cc.p54 * X3[kc] - Cc.p22 * X4[kc]) /
                    p.dzeta;
    }
    {
        const long ia = c + sx, jb = c + sy, kc = c + 1;
        const SC A = nb_split(U_, J, met, ia, 0, K, e);
        const SC B = nb_split(U_, J, met, jb, 1, K, e);
        const SC Cc = nb_split(U_, J, met, kc, 2, K, e);
        UU[0] = (+A.m11 * X0[ia] + A.p12 * X1[ia] + A.p13 * X2[ia] + A.p14 * X3[ia]) / p.dxi +
                (+B.m11 * X0[jb] + B.p12 * X1[jb] + B.p13 * X2[jb] + B.p14 * X3[jb]) / p.deta +
                (+Cc.m11 * X0[kc] + Cc.p12 * X1[kc] + Cc.p13 * X2[kc] + Cc.p14 * X3[kc]) / p.dzeta;
        UU[1] = (+A.p21 * X0[ia] + A.m22 * X1[ia]) / p.dxi + (+B.p21 * X0[jb] + B.m22 * X1[jb]) / p.deta +
                (+Cc.p21 * X0[kc] + Cc.m22 * X1[kc]) / p.dzeta;
        UU[2] = (+A.p31 * X0[ia] + A.m22 * X2[ia]) / p.dxi + (+B.p31 * X0[jb] + B.m22 * X2[jb]) / p.deta +
                (+Cc.p31 * X0[kc] + Cc.m22 * X2[kc]) / p.dzeta;
        UU[3] = (+A.p41 * X0[ia] + A.m22 * X3[ia]) / p.dxi + (+B.p41 * X0[jb] + B.m22 * X3[jb]) / p.deta +
                (+Cc.p41 * X0[kc] + Cc.m22 * X3[kc]) / p.dzeta;
        UU[4] = (+A.p51 * X0[ia] + A.p52 * X1[ia] + A.p53 * X2[ia] + A.p54 * X3[ia] + A.m22 * X4[ia]) / p.dxi +
                (+B.p51 * X0[jb] + B.p52 * X1[jb] + B.p53 * X2[jb] + B.p54 * X3[jb] + B.m22 * X4[jb]) / p.deta +
                (+Cc.p51 * X0[kc] + Cc.p52 * X1[kc] + Cc.p53 * X2[kc] + Cc.p54 * X3[kc] + Cc.m22 * X4[kc]) /
                    p.dzeta;
    }
    const CP q = cell_prim(U_, J, c, K, e);
    double m[9];
    load9(met, c, m);
    const DG D = diag(q, m, p, deltaT);
    NRn.v[0][c] = D.d11 * (MR.v[0][c] - LL[0] - UU[0]);
    NRn.v[1][c] = D.d22 * (MR.v[1][c] - LL[1] - UU[1]);
    NRn.v[2][c] = D.d22 * (MR.v[2][c] - LL[2] - UU[2]);
    NRn.v[3][c] = D.d22 * (MR.v[3][c] - LL[3] - UU[3]);
    NRn.v[4][c] = D.d51 * (MR.v[0][c] - LL[0] - UU[0]) + D.d22 * (MR.v[4][c] - LL[4] - UU[4]);
}

void launch_dplur_sweep(const Geo &g, const Phys &p, double deltaT, double e, C5 U_, const double *J, M9 met,
                        C5 MR, C5 NLold, F5 NRnew, hipStream_t s)
{
    dim3 b(64, 4);
    dim3 gr((g.NZ - 1 + b.x - 1) / b.x, (g.NY - 1 + b.y - 1) / b.y, g.nx);
    hipLaunchKernelGGL(k_dplur, gr, b, 0, s, g, p, deltaT, e, U_, J, met, MR, NLold, NRnew);
}

__global__ void k_copy_interior(Geo g, C5 src, F5 dst)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; v++) dst.v[v][c] = src.v[v][c];
}

void launch_copy_interior(const Geo &g, C5 src, F5 dst, hipStream_t s)
{
    dim3 b(64, 4);
    dim3 gr((g.NZ - 1 + b.x - 1) / b.x, (g.NY - 1 + b.y - 1) / b.y, g.nx);
    hipLaunchKernelGGL(k_copy_interior, gr, b, 0, s, g, src, dst);
}

// ------------------------------------------------------ update + norm
// Dual_time_stepping.cpp:1702-1743 (DPLUSGS.cpp:1163-1202 measures the
// pressure increment and squares again).  Thread t covers interior cell t
// in (i, j, k) order; each block writes 5 partial sums.
__global__ void __launch_bounds__(256) k_update_norm(Geo g, Phys p, F5 U_, const double *__restrict__ J, C5 NR,
                                                     F5 Res, int write_res, int dplusgs, double *partials)
{
    const long ncell = (long)g.nx * g.ny * g.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double x[5] = {0, 0, 0, 0, 0};
    if (t < ncell) {
        const int k = 2 + (int)(t % g.nz);
        const int j = 2 + (int)((t / g.nz) % g.ny);
        const int i = 3 + (int)(t / ((long)g.nz * g.ny));
        const long c = ix(g, i, j, k);
        const double K = p.K;
        const double Jc = J[c];
        const double U1 = U_.v[0][c];
        const double rhoold = U1 * Jc;
        const double Uold = U_.v[1][c] / U1;
        const double Vold = U_.v[2][c] / U1;
        const double Wold = U_.v[3][c] / U1;
        const double VVold = Uold * Uold + Vold * Vold + Wold * Wold;
        const double Pold = (U_.v[4][c] * Jc - 0.5 * rhoold * VVold) * (K - 1);
        const double Told = Pold / rhoold;
        const double P = Pold + NR.v[0][c] * Jc;
        const double Uu = Uold + NR.v[1][c] * Jc;
        const double V = Vold + NR.v[2][c] * Jc;
        const double W = Wold + NR.v[3][c] * Jc;
        const double T = Told + NR.v[4][c] * Jc;
        const double rho = P / T;
        U_.v[0][c] = rho / Jc;
        U_.v[1][c] = rho * Uu / Jc;
        U_.v[2][c] = rho * V / Jc;
        U_.v[3][c] = rho * W / Jc;
        U_.v[4][c] = (P / (K - 1) + 0.5 * rho * (Uu * Uu + V * V + W * W)) / Jc;
        x[0] = dplusgs ? (P - Pold) * (P - Pold) : (rho - rhoold) * (rho - rhoold);
        x[1] = (Uu - Uold) * (Uu - Uold);
        x[2] = (V - Vold) * (V - Vold);
        x[3] = (W - Wold) * (W - Wold);
        x[4] = (T - Told) * (T - Told);
        if (write_res) {
#pragma unroll
            for (int v = 0; v < 5; v++) Res.v[v][c] = x[v];
        }
        if (dplusgs) {
#pragma unroll
            for (int v = 0; v < 5; v++) x[v] = x[v] * x[v];
        }
    }
    double o[5];
    block_sum<5>(x, o);
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 5; v++) partials[5 * (long)blockIdx.x + v] = o[v];
}

int launch_update_norm(const Geo &g, const Phys &p, F5 U_, const double *J, C5 NR, F5 Res, int write_res,
                       int dplusgs, double *partials, hipStream_t s)
{
    const long ncell = (long)g.nx * g.ny * g.nz;
    const int nb = (int)((ncell + 255) / 256);
    hipLaunchKernelGGL(k_update_norm, dim3(nb), dim3(256), 0, s, g, p, U_, J, NR, Res, write_res, dplusgs,
                       partials);
    return nb;
}

// one block: each thread sums a contiguous run of partials in order, then
// the block combines the runs in thread order
// er may be pinned host memory (written through its device address); with
// counters, the wavefront's ticket and error flag are copied to flags too
__global__ void __launch_bounds__(1024) k_finish_norm(int nb, const double *__restrict__ partials, double DN,
                                                      double *er, const int *counters, int *flags)
{
    // thread t sums partials t, t + 1024, ... in that order (coalesced),
    // then the block sum in its fixed order
    double x[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int v = 0; v < 5; v++) x[v] += partials[5 * (long)b + v];
    double o[5];
    block_sum<5>(x, o);
    if (threadIdx.x == 0) {
        // system scope: the host reads these after the kernel's completion
        // event, whatever caching the mapping of its memory allows
#pragma unroll
        for (int v = 0; v < 5; v++)
            __hip_atomic_store(reinterpret_cast<unsigned long long *>(er + v),
                               (unsigned long long)__double_as_longlong(sqrt(o[v]) * DN), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        if (counters) {
            __hip_atomic_store(flags, counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(flags + 1, counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

void launch_finish_norm(int nblocks, const double *partials, double DN, int squared_twice, double *er_dev,
                        hipStream_t s, const int *counters, int *flags)
{
    (void)squared_twice;
    hipLaunchKernelGGL(k_finish_norm, dim3(1), dim3(1024), 0, s, nblocks, partials, DN, er_dev, counters, flags);
}


// =================================================================
// Wavefront LU-SGS.
//
// Pass 1 (k_split_coef, fully parallel): per cell 21 pairs of the cell's
// split coefficients along x, y and z (14 values each, already divided by
// the grid spacing of that direction), stored as CT (double, or float for
// the mixed-precision sweep), and for interior cells the diagonal (d11,
// d22, d51) in two FP64 pairs of a separate array.  Cells are grouped in
// blocks of CB consecutive k (one tile row); inside a block pair p of all
// CB cells is contiguous, so a tile row reads one pair with a single load
// per lane over contiguous bytes, and all data of a cell's block lies
// within a few kilobytes.  The sweep is only the preconditioner of the
// inner iteration: an FP32 split (with the FP64 diagonal) changes the
// iteration's convergence path, not the state it converges to.
//
// Pass 2 (k_lusgs_wave, one launch per sweep): the (j,k) plane is cut into
// TJ x TK tiles; a workgroup owns one tile, one thread per (j,k) column,
// and marches i.  At local step s thread (jl,kl) updates i = s - jl - kl,
// so its i-1, j-1 and k-1 neighbours were updated at step s-1: inside the
// tile they come from LDS, across tile edges (and from the ghost planes)
// from an exchange buffer.  Before the sweep every interior exchange entry
// holds a sentinel NaN; a tile-edge thread writes its result with an
// agent-scope atomic store and the consumer spins on an atomic load until
// the sentinel is gone, so each value is validated on its own and no
// ordering between locations is needed.  Tiles are claimed through a
// ticket in topological order (anti-diagonals), so a workgroup only ever
// waits on tiles held by workgroups that are already running.  Spins are
// bounded; an expired spin sets *err and the kernel still runs to
// completion.  The coefficients of the next cell of each column are loaded
// one step ahead into one of two register sets used alternately.
//
// Compared with the plane-by-plane kernels (which follow the reference's
// expressions literally) the only arithmetic difference is that the
// division by dxi/deta/dzeta is applied to each coefficient instead of to
// each partial sum: results agree to rounding (a few ulp).
// =================================================================
constexpr int NSPLIT = 21;             // 3 x 7 split pairs per cell
constexpr int CB = WAVE_TK;            // cells per block along k
constexpr int CBLK = NSPLIT * CB * 2;  // split elements per block
constexpr int DBLK = 2 * CB * 2;       // diagonal doubles per block: (d11, d22), (d51, pad)
constexpr int CKOFF = CB - 2;          // the first interior k starts a block
constexpr long CPAIR = 2 * CB;         // elements from one pair to the next

// blocks per (i,j) row
__host__ __device__ inline long coef_nkb(const Geo &g) { return (g.Zm + CKOFF + CB - 1) / CB; }

// offset (in elements) of pair 0 of cell (i,j,k) in a blocked array of
// `blk` elements per block
__device__ __forceinline__ long block_at(const Geo &g, long nkb, long blk, int i, int j, int k)
{
    const int kk = k + CKOFF;
    return (((long)i * g.Ym + j) * nkb + kk / CB) * blk + (kk % CB) * 2;
}

template <typename CT> struct Pair2;
template <> struct Pair2<double> {
    typedef double2 t;
    static __device__ __forceinline__ double2 make(double a, double b) { return make_double2(a, b); }
};
template <> struct Pair2<float> {
    typedef float2 t;
    static __device__ __forceinline__ float2 make(double a, double b) { return make_float2((float)a, (float)b); }
};
// slot order of one direction: the ten values both sweeps use, then the
// forward pair (p11, p22), then the backward pair (m11, m22)
enum { S_P12, S_P13, S_P14, S_P21, S_P31, S_P41, S_P51, S_P52, S_P53, S_P54, S_P11, S_P22, S_M11, S_M22 };
constexpr unsigned long long SENTINEL = 0x7FF4DEADBEEF0001ULL;

template <typename CT>
__global__ void __launch_bounds__(256) k_split_coef(Geo g, Phys p, double deltaT, double e, C5 U_,
                                                    const double *__restrict__ J, M9 met, CT *__restrict__ coef,
                                                    double *__restrict__ dgn)
{
    // box i=2..gend+1, j=1..NYY, k=1..NZZ
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 2 + blockIdx.z;
    if (k > g.NZZ || j > g.NYY) return;
    const long c = ix(g, i, j, k);
    typedef typename Pair2<CT>::t V2;
    const long nkb = coef_nkb(g);
    V2 *o = reinterpret_cast<V2 *>(coef + block_at(g, nkb, CBLK, i, j, k));
    const CP q = cell_prim(U_, J, c, p.K, e);
    const double dd[3] = {p.dxi, p.deta, p.dzeta};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const SC s = split(q, met.m[3 * d][c], met.m[3 * d + 1][c], met.m[3 * d + 2][c], p.K, d != 0);
        const double h = dd[d];
        V2 *b = o + 7 * d * CB;
        b[0 * CB] = Pair2<CT>::make(s.p12 / h, s.p13 / h);
        b[1 * CB] = Pair2<CT>::make(s.p14 / h, s.p21 / h);
        b[2 * CB] = Pair2<CT>::make(s.p31 / h, s.p41 / h);
        b[3 * CB] = Pair2<CT>::make(s.p51 / h, s.p52 / h);
        b[4 * CB] = Pair2<CT>::make(s.p53 / h, s.p54 / h);
        b[5 * CB] = Pair2<CT>::make(s.p11 / h, s.p22 / h);
        b[6 * CB] = Pair2<CT>::make(s.m11 / h, s.m22 / h);
    }
    if (i >= 3 && i <= g.gend && j >= 2 && j <= g.NY && k >= 2 && k <= g.NZ) {
        double m[9];
        load9(met, c, m);
        const DG D = diag(q, m, p, deltaT);
        double2 *od = reinterpret_cast<double2 *>(dgn + block_at(g, nkb, DBLK, i, j, k));
        od[0] = make_double2(D.d11, D.d22);
        od[CB] = make_double2(D.d51, 0.0);
    }
}

// the sentinel into the exchange entry of every interior cell
__global__ void k_fill_sentinel(Geo g, double *__restrict__ xbuf)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; v++) xbuf[5 * c + v] = __longlong_as_double((long long)SENTINEL);
}

// The sentinel into the exchange entries one march sweep polls, and the
// sweep's tile ticket to zero.  A tile's consumers read the row of cells
// just across its j edge and the column just across its k edge, which are
// the rows (j-2) % TJ == rj and columns (k-2) % TK == rk of every x plane
// (rj = TJ-1, rk = TK-1 forward; 0, 0 backward); no other entry is read.
// blockIdx.y 0: the rows, 5 nz consecutive doubles each; 1: the columns.
// The rows are polled in xrow, the columns in xcol (one buffer for both when
// the increments themselves are exchanged, two when the j and k products
// differ).
__global__ void k_fill_sentinel_edges(Geo g, double *__restrict__ xrow, double *__restrict__ xcol, int TJ, int TK,
                                      int rj, int rk, int *__restrict__ ticket)
{
    const int i = 3 + blockIdx.z;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ticket && t == 0 && blockIdx.y == 0 && blockIdx.z == 0) *ticket = 0;
    const double S = __longlong_as_double((long long)SENTINEL);
    const int nrow = rj < g.ny ? (g.ny - 1 - rj) / TJ + 1 : 0;
    const int ncol = rk < g.nz ? (g.nz - 1 - rk) / TK + 1 : 0;
    if (blockIdx.y == 0) {
        const long per = 5L * g.nz;
        if (t >= per * nrow) return;
        const int j = 2 + rj + (int)(t / per) * TJ;
        xrow[5 * ix(g, i, j, 2) + t % per] = S;
    } else {
        if (t >= 5L * g.ny * ncol) return;
        const long cell = t / 5;
        const int j = 2 + (int)(cell / ncol), k = 2 + rk + (int)(cell % ncol) * TK;
        xcol[5 * ix(g, i, j, k) + t % 5] = S;
    }
}

__global__ void k_reset_ints(int *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// one direction's 12 values for a sweep: P12..P54, then g11, g22 (the
// p- or m-pair)
template <typename CT>
struct Dir12 { CT v[12]; };

template <bool FWD, typename CT>
__device__ __forceinline__ void load_dir(Dir12<CT> &s, const CT *__restrict__ r)
{
    typedef typename Pair2<CT>::t V2;
    const V2 *b = reinterpret_cast<const V2 *>(r);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const V2 t = b[q * CB];
        s.v[2 * q] = t.x;
        s.v[2 * q + 1] = t.y;
    }
    const V2 t = b[(FWD ? 5 : 6) * CB];
    s.v[10] = t.x;
    s.v[11] = t.y;
}

template <typename CT>
struct Pref { Dir12<CT> A, B, C; double d11, d22, d51; double r[5]; };

// rec: the own cell's split record; oA/oB/oC: offsets of the upstream
// neighbours' records along i, j, k; drec: the own cell's diagonal
template <bool FWD, typename CT>
__device__ __forceinline__ void load_pref(Pref<CT> &P, const CT *__restrict__ rec, long oA, long oB, long oC,
                                          const double *__restrict__ drec, const C5 &R, long c)
{
    load_dir<FWD>(P.A, rec + oA);
    load_dir<FWD>(P.B, rec + oB + 7 * CPAIR);
    load_dir<FWD>(P.C, rec + oC + 14 * CPAIR);
    const double2 *b = reinterpret_cast<const double2 *>(drec);
    const double2 t0 = b[0], t1 = b[CB];
    P.d11 = t0.x;
    P.d22 = t0.y;
    P.d51 = t1.x;
#pragma unroll
    for (int v = 0; v < 5; v++) P.r[v] = R.v[v][c];
}

__device__ __forceinline__ unsigned long long aload(const double *p)
{
    return __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// issue the five exchange loads of one neighbour without waiting (the
// exchange buffer holds the five values of a cell together: xbuf[5c + v])
__device__ __forceinline__ void issue5(const double *xbuf, long N, long c, unsigned long long b[5])
{
#pragma unroll
    for (int v = 0; v < 5; v++) b[v] = aload(xbuf + 5 * c + v);
}

// where `need`, re-poll the entries that still hold the sentinel (bounded)
// and take them as the neighbour's values
__device__ __forceinline__ void settle5(const double *xbuf, long N, long c, bool need, unsigned long long b[5],
                                        double out[5], int max_spin, int *err)
{
#pragma unroll
    for (int v = 0; v < 5; v++) {
        int n = 0;
        while (need && b[v] == SENTINEL && n < max_spin) {
            __builtin_amdgcn_s_sleep(1);
            b[v] = aload(xbuf + 5 * c + v);
            n++;
        }
        if (need && b[v] == SENTINEL) atomicOr(err, 1);
        if (need) out[v] = __longlong_as_double((long long)b[v]);
    }
}

// per-lane constants of the march down one (j,k) column
struct Lane {
    int tid, jl, kl;
    bool col, edge_out;
    long c0;                // cell (i = 0, j, k), clamped into the domain
    long r0, d0;            // split and diagonal offsets of the column's first cell
    long tline;             // this wave's line of the trash buffer
};

template <typename CT>
struct WaveArgs {
    Geo g;
    const CT *coef;
    const double *dgn;
    C5 R;
    F5 X;
    double *xbuf, *trash;
    int *err;
    int max_spin;
    long sxs, oA, oB, oC, d, dcol, ddcol;   // dcol/ddcol: split/diagonal offset of the next cell of a column
};

// local step il clamped to the column, so every lane has a valid address
// at every step
template <typename CT>
__device__ __forceinline__ int icl(const WaveArgs<CT> &w, int il)
{
    return il < 0 ? 0 : (il >= w.g.nx ? w.g.nx - 1 : il);
}

// cell of local step il
template <bool FWD, typename CT>
__device__ __forceinline__ long cell_at(const WaveArgs<CT> &w, const Lane &L, int il)
{
    const int c = icl(w, il);
    return L.c0 + (long)(FWD ? 3 + c : w.g.gend - c) * w.sxs;
}

// FWD: X = NL from R = MR, NL = D (MR - L NL).  BWD: X = NR from R = NL,
// NR = NL - D U NR.  One step of the march.  Coefficient loads are issued
// by every lane (clamped addresses) and lanes with nothing to do store X
// into a trash line, so no branch separates a prefetch from its use; the
// tile-edge loads and stores are confined to the lanes that need them.  PAR is the step's parity (which LDS
// half is written); Pc/nbc/ncc hold the values loaded for this step,
// Pn/nbn/ncn receive the next step's.
template <int TJ, int TK, bool FWD, int PAR, typename CT>
__device__ __forceinline__ void wave_step(const WaveArgs<CT> &w, const Lane &L, int s, double prev[5],
                                          double (*sx)[5][TJ * TK], Pref<CT> &Pc, Pref<CT> &Pn,
                                          unsigned long long *nbc, unsigned long long *nbn, unsigned long long *ncc,
                                          unsigned long long *ncn)
{
    const int il = s - L.jl - L.kl;
    const bool act = L.col && il >= 0 && il < w.g.nx;
    __syncthreads();
    const long c = cell_at<FWD>(w, L, il);
    double xb[5], xc[5];
    const int tb = L.tid >= TK ? L.tid - TK : L.tid, tc = L.tid >= 1 ? L.tid - 1 : L.tid;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        xb[v] = sx[PAR ^ 1][v][tb];
        xc[v] = sx[PAR ^ 1][v][tc];
    }
    settle5(w.xbuf, w.g.N, c + w.d * w.g.Zm, act && L.jl == 0, nbc, xb, w.max_spin, w.err);
    settle5(w.xbuf, w.g.N, c + w.d, act && L.kl == 0, ncc, xc, w.max_spin, w.err);
    {   // prefetch for the next step: the tile-edge values first (only the
        // lanes that consume them), then the coefficients of every lane
        const long cn = cell_at<FWD>(w, L, il + 1);
        if (L.jl == 0) issue5(w.xbuf, w.g.N, cn + w.d * w.g.Zm, nbn);
        if (L.kl == 0) issue5(w.xbuf, w.g.N, cn + w.d, ncn);
        const long ic = icl(w, il + 1);
        load_pref<FWD>(Pn, w.coef + L.r0 + ic * w.dcol, w.oA, w.oB, w.oC, w.dgn + L.d0 + ic * w.ddcol, w.R, cn);
    }
    const CT *A = Pc.A.v, *B = Pc.B.v, *Cc = Pc.C.v;
    // the off-diagonal products in the coefficients' precision (the FP64
    // path converts nothing)
    CT xa[5], xbc[5], xcc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        xa[v] = (CT)prev[v];
        xbc[v] = (CT)xb[v];
        xcc[v] = (CT)xc[v];
    }
    double out[5];
    if (FWD) {
        const CT LL0 = (-A[10] * xa[0] - A[0] * xa[1] - A[1] * xa[2] - A[2] * xa[3]) +
                           (-B[10] * xbc[0] - B[0] * xbc[1] - B[1] * xbc[2] - B[2] * xbc[3]) +
                           (-Cc[10] * xcc[0] - Cc[0] * xcc[1] - Cc[1] * xcc[2] - Cc[2] * xcc[3]);
        const CT LL1 = (-A[3] * xa[0] - A[11] * xa[1]) + (-B[3] * xbc[0] - B[11] * xbc[1]) +
                           (-Cc[3] * xcc[0] - Cc[11] * xcc[1]);
        const CT LL2 = (-A[4] * xa[0] - A[11] * xa[2]) + (-B[4] * xbc[0] - B[11] * xbc[2]) +
                           (-Cc[4] * xcc[0] - Cc[11] * xcc[2]);
        const CT LL3 = (-A[5] * xa[0] - A[11] * xa[3]) + (-B[5] * xbc[0] - B[11] * xbc[3]) +
                           (-Cc[5] * xcc[0] - Cc[11] * xcc[3]);
        const CT LL4 = (-A[6] * xa[0] - A[7] * xa[1] - A[8] * xa[2] - A[9] * xa[3] - A[11] * xa[4]) +
                           (-B[6] * xbc[0] - B[7] * xbc[1] - B[8] * xbc[2] - B[9] * xbc[3] - B[11] * xbc[4]) +
                           (-Cc[6] * xcc[0] - Cc[7] * xcc[1] - Cc[8] * xcc[2] - Cc[9] * xcc[3] - Cc[11] * xcc[4]);
        out[0] = Pc.d11 * (Pc.r[0] - (double)LL0);
        out[1] = Pc.d22 * (Pc.r[1] - (double)LL1);
        out[2] = Pc.d22 * (Pc.r[2] - (double)LL2);
        out[3] = Pc.d22 * (Pc.r[3] - (double)LL3);
        out[4] = Pc.d51 * (Pc.r[0] - (double)LL0) + Pc.d22 * (Pc.r[4] - (double)LL4);
    } else {
        const CT UU0 = (+A[10] * xa[0] + A[0] * xa[1] + A[1] * xa[2] + A[2] * xa[3]) +
                           (+B[10] * xbc[0] + B[0] * xbc[1] + B[1] * xbc[2] + B[2] * xbc[3]) +
                           (+Cc[10] * xcc[0] + Cc[0] * xcc[1] + Cc[1] * xcc[2] + Cc[2] * xcc[3]);
        const CT UU1 = (+A[3] * xa[0] + A[11] * xa[1]) + (+B[3] * xbc[0] + B[11] * xbc[1]) +
                           (+Cc[3] * xcc[0] + Cc[11] * xcc[1]);
        const CT UU2 = (+A[4] * xa[0] + A[11] * xa[2]) + (+B[4] * xbc[0] + B[11] * xbc[2]) +
                           (+Cc[4] * xcc[0] + Cc[11] * xcc[2]);
        const CT UU3 = (+A[5] * xa[0] + A[11] * xa[3]) + (+B[5] * xbc[0] + B[11] * xbc[3]) +
                           (+Cc[5] * xcc[0] + Cc[11] * xcc[3]);
        const CT UU4 = (+A[6] * xa[0] + A[7] * xa[1] + A[8] * xa[2] + A[9] * xa[3] + A[11] * xa[4]) +
                           (+B[6] * xbc[0] + B[7] * xbc[1] + B[8] * xbc[2] + B[9] * xbc[3] + B[11] * xbc[4]) +
                           (+Cc[6] * xcc[0] + Cc[7] * xcc[1] + Cc[8] * xcc[2] + Cc[9] * xcc[3] + Cc[11] * xcc[4]);
        out[0] = Pc.r[0] - Pc.d11 * (double)UU0;
        out[1] = Pc.r[1] - Pc.d22 * (double)UU1;
        out[2] = Pc.r[2] - Pc.d22 * (double)UU2;
        out[3] = Pc.r[3] - Pc.d22 * (double)UU3;
        out[4] = Pc.r[4] - (Pc.d51 * (double)UU0 + Pc.d22 * (double)UU4);
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double *px = act ? w.X.v[v] + c : w.trash + L.tline + v;
        *px = out[v];
        sx[PAR][v][L.tid] = out[v];
        prev[v] = act ? out[v] : prev[v];
    }
    if (act && L.edge_out) {
#pragma unroll
        for (int v = 0; v < 5; v++)
            __hip_atomic_store(reinterpret_cast<unsigned long long *>(w.xbuf + 5 * c + v),
                               (unsigned long long)__double_as_longlong(out[v]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// trash: 256 doubles per tile of scratch that no one reads; lanes of one
// wave with nothing to store all write the same 128-byte line, so a store
// instruction costs one extra line and waves never share one
template <int TJ, int TK, bool FWD, typename CT>
__global__ void __launch_bounds__(TJ *TK) k_lusgs_wave(Geo g, const CT *__restrict__ coef,
                                                     const double *__restrict__ dgn, C5 R, F5 X, double *xbuf,
                                                     double *trash, const int *__restrict__ order, int *ticket,
                                                     int *err, int max_spin)
{
    static_assert((TJ * TK + 63) / 64 <= 16, "one trash line per wave, at most 16 waves");
    __shared__ int s_t;
    __shared__ double sx[2][5][TJ * TK];
    Lane L;
    L.tid = threadIdx.x;
    L.kl = L.tid % TK;
    L.jl = L.tid / TK;
    if (L.tid == 0) s_t = atomicAdd(ticket, 1);
    __syncthreads();
    const int packed = order[s_t];
    const int TJi = packed >> 16, TKi = packed & 0xffff;
    const int jo = TJi * TJ + L.jl, ko = TKi * TK + L.kl;
    L.col = (jo < g.ny) && (ko < g.nz);
    const int joc = jo < g.ny ? jo : g.ny - 1, koc = ko < g.nz ? ko : g.nz - 1;
    const int j = FWD ? 2 + joc : g.NY - joc;
    const int k = FWD ? 2 + koc : g.NZ - koc;
    L.c0 = ix(g, 0, j, k);
    L.edge_out = (L.jl == TJ - 1) || (L.kl == TK - 1);
    L.tline = ((long)blockIdx.x * 16 + L.tid / 64) * 16;
    WaveArgs<CT> w;
    w.g = g;
    w.coef = coef;
    w.dgn = dgn;
    w.R = R;
    w.X = X;
    w.xbuf = xbuf;
    w.trash = trash;
    w.err = err;
    w.max_spin = max_spin;
    w.sxs = (long)g.Ym * g.Zm;
    w.d = FWD ? -1 : 1;   // towards the upstream neighbours
    {
        const long nkb = coef_nkb(g);
        const int i0 = FWD ? 3 : g.gend;
        w.dcol = -w.d * g.Ym * nkb * CBLK;
        w.ddcol = -w.d * g.Ym * nkb * DBLK;
        w.oA = -w.dcol;
        w.oB = w.d * nkb * CBLK;
        L.r0 = block_at(g, nkb, CBLK, i0, j, k);
        L.d0 = block_at(g, nkb, DBLK, i0, j, k);
        w.oC = block_at(g, nkb, CBLK, i0, j, k + (int)w.d) - L.r0;
    }
    const int S = g.nx + TJ + TK - 2;
    double prev[5];
    Pref<CT> P0, P1;
    unsigned long long nb0[5], nb1[5], nc0[5], nc1[5];
    {
        const long c = cell_at<FWD>(w, L, 0);
#pragma unroll
        for (int v = 0; v < 5; v++) prev[v] = X.v[v][c + w.d * w.sxs];   // ghost plane (zero)
        const long cs = cell_at<FWD>(w, L, -(L.jl + L.kl));               // the cell of step 0
        if (L.jl == 0) issue5(xbuf, g.N, cs + w.d * g.Zm, nb0);
        if (L.kl == 0) issue5(xbuf, g.N, cs + w.d, nc0);
        const long ic = icl(w, -(L.jl + L.kl));
        load_pref<FWD>(P0, coef + L.r0 + ic * w.dcol, w.oA, w.oB, w.oC, dgn + L.d0 + ic * w.ddcol, R, cs);
    }
    for (int s = 0; s < S; s += 2) {
        wave_step<TJ, TK, FWD, 0>(w, L, s, prev, sx, P0, P1, nb0, nb1, nc0, nc1);
        if (s + 1 < S) wave_step<TJ, TK, FWD, 1>(w, L, s + 1, prev, sx, P1, P0, nb1, nb0, nc1, nc0);
    }
}

// ---------------------------------------------------------------- march layout
// The tiled wavefront with its coefficients stored in the order the march
// reads them.  Lane (jl, kl) of tile T works at step s on the cell whose
// forward march index t = (i-3) + jl + kl equals s, so the 64 cells of a
// tile at one step lie on a diagonal of (i, j, k): in any i-major layout
// their records are 64 scattered pieces.  Here the record of the cell at
// march position (T, t, lane) is stored in row T*S + t of the records, pair
// p of it at [row][p][lane], so a tile reads one contiguous block per step.
// The backward sweep runs the same schedule in reverse (t = S-1-s, tiles in
// descending anti-diagonals), which visits the cells in a reverse
// topological order, so it reads the same records.  Records are written for
// interior cells only: where a neighbour is a ghost cell its increment is
// zero and the cell's own record stands in for the ghost's coefficients.
// The sweep's right-hand sides and increments are held in the same order
// ([row][variable][lane]): the split kernel gathers MR into it, the forward
// sweep writes NL there for the backward sweep, and a parallel kernel
// scatters NL and NR back into the field layout afterwards.  A wave's loads
// and stores are then contiguous blocks; the scattered accesses happen in
// kernels that run every cell at once, where they cost bandwidth and not
// latency on the dependent march.
//
// Every array has MPAD rows of padding before and after, so the rows a lane
// reads at steps where it has no cell (its neighbours' rows shifted by up
// to TJ+TK steps) are inside the allocation without clamping.  Addresses
// are a wave-uniform row base plus a per-lane 32-bit offset fixed for the
// whole sweep.
struct March { int ntj, ntk, S; };

constexpr int MPAD = MARCH_TJ + MARCH_TK + 2;

// padded row of march position (T, t)
__host__ __device__ inline long march_row(const March &M, int T, int t) { return (long)T * M.S + t + MPAD; }

static March march_of(const Geo &g)
{
    March M;
    M.ntj = (g.ny + MARCH_TJ - 1) / MARCH_TJ;
    M.ntk = (g.nz + MARCH_TK - 1) / MARCH_TK;
    M.S = g.nx + MARCH_TJ + MARCH_TK - 2;
    return M;
}

static long march_rows(const Geo &g)
{
    const March M = march_of(g);
    return (long)M.ntj * M.ntk * M.S + 2 * MPAD;
}
long march_coef_pairs(const Geo &g) { return march_rows(g) * NSPLIT * 64; }
long march_diag_pairs(const Geo &g) { return march_rows(g) * 2 * 64; }
long march_field_doubles(const Geo &g) { return march_rows(g) * 5 * 64; }

// the per-lane record offsets of a sweep span ntk tiles of rows: they must
// fit 32 bits
bool march_fits(const Geo &g, int fp32)
{
    const March M = march_of(g);
    const double rowb = (double)NSPLIT * 64 * 2 * (fp32 ? 4 : 8);
    return ((double)(M.ntk + 1) * M.S + 4.0 * MPAD) * rowb < 4.0e9;
}

// Gather and scatter between the field layout and march order: one wave
// per tile column row jl and block of MR_ROWS = 64/TK march steps; its 64
// threads cover the positions (step t0+r, lane jl*TK+kl), ordered by the
// cell's plane a = r - kl (then kl), so that the field-layout side of each
// memory instruction is a few runs of consecutive k (one per plane) and the
// march side is MR_ROWS runs of TK consecutive lanes (one per step).
constexpr int MR_ROWS = 64 / MARCH_TK;
static_assert(MR_ROWS * MARCH_TK == 64, "a wave covers whole steps of one tile row");

struct MPos { int r, kl; };
__device__ __forceinline__ MPos mpos_of_lane(int q)
{
    constexpr int R = MR_ROWS, TK = MARCH_TK;
    // planes a = r - kl + TK - 1 in ascending order, kl ascending inside
    for (int a = 0; a < R + TK - 1; a++) {
        // kl range of plane a: r = a - (TK - 1) + kl in [0, R)
        const int lo = (TK - 1) - a > 0 ? (TK - 1) - a : 0;
        const int hi = (R - 1) + (TK - 1) - a < TK - 1 ? (R - 1) + (TK - 1) - a : TK - 1;
        const int n = hi - lo + 1;
        if (q < n) return MPos{a - (TK - 1) + lo + q, lo + q};
        q -= n;
    }
    return MPos{0, 0};
}

// Work units: (MG_KP adjacent tile columns, 4 / MG_KP tile rows, block of
// MR_ROWS steps, tile row TJi); a block of 4 waves takes one, a wave per
// (column, row): the waves of a row cover 8 MG_KP consecutive k, whole cache
// lines of a field-layout row (4 columns: 32 k, 2.6 % faster gather and
// update than 2).  The grid is XCD-aware: workgroup b runs on XCD b % 8,
// and XCD x takes the units of a contiguous group of tile columns,
// k-neighbours first, then step blocks.  The planes two step blocks share
// are then read through one XCD's L2.  (With another placement of workgroups
// the result is the same, only slower.)
constexpr int MG_XCD = 8;
constexpr int MG_KP = 4;   // tile columns per block
struct MGather {
    int T, t, l, jl, kl, jo, ko, il;
    bool ok;
};
__device__ __forceinline__ MGather mgather_at(const Geo &g, const March &M)
{
    constexpr int TJ = MARCH_TJ, TK = MARCH_TK, WPB = 4, JW = WPB / MG_KP;
    constexpr int JB = (TJ + JW - 1) / JW;   // blocks per tile row set
    MGather q;
    const int wv = threadIdx.x >> 6;
    const int x = blockIdx.x % MG_XCD, u = blockIdx.x / MG_XCD;
    const int k0 = x * M.ntk / MG_XCD, nk = (x + 1) * M.ntk / MG_XCD - k0;
    const int nkp = (nk + MG_KP - 1) / MG_KP;
    const int ntb = (M.S + MR_ROWS - 1) / MR_ROWS;
    q.ok = false;
    if (nk == 0 || u >= nkp * JB * ntb * M.ntj) return q;
    const int kt = (u % nkp) * MG_KP + wv % MG_KP;   // the wave's column in the XCD's group
    const int TKi = k0 + kt;
    int r = u / nkp;
    const int jb = r % JB;
    r /= JB;
    const int tb = r % ntb, TJi = r / ntb;
    q.jl = jb * JW + wv / MG_KP;
    const MPos P = mpos_of_lane(threadIdx.x & 63);
    q.T = TJi * M.ntk + TKi;
    q.t = tb * MR_ROWS + P.r;
    q.kl = P.kl;
    q.l = q.jl * TK + q.kl;
    q.jo = TJi * TJ + q.jl;
    q.ko = TKi * TK + q.kl;
    q.il = q.t - q.jl - q.kl;
    q.ok = kt < nk && q.jl < TJ && q.t < M.S && q.il >= 0 && q.il < g.nx && q.jo < g.ny && q.ko < g.nz;
    return q;
}
static dim3 mgather_grid(const March &M)
{
    constexpr int JB = (MARCH_TJ + 4 / MG_KP - 1) / (4 / MG_KP);
    const long cols = (M.ntk + MG_XCD - 1) / MG_XCD;
    const long per_xcd = (cols + MG_KP - 1) / MG_KP * JB * ((M.S + MR_ROWS - 1) / MR_ROWS) * M.ntj;
    return dim3((unsigned)(MG_XCD * per_xcd));
}

template <typename CT>
__global__ void __launch_bounds__(256) k_split_march(Geo g, Phys p, double deltaT, double e, C5 U_,
                                                     const double *__restrict__ J, M9 met, C5 MR, March M,
                                                     CT *__restrict__ coefM, double *__restrict__ dgnM,
                                                     double *__restrict__ mrM)
{
    const MGather q0 = mgather_at(g, M);
    if (!q0.ok) return;
    const int T = q0.T, t = q0.t, l = q0.l, jo = q0.jo, ko = q0.ko, il = q0.il;
    const long c = ix(g, 3 + il, 2 + jo, 2 + ko);
    const long row = march_row(M, T, t);
    typedef typename Pair2<CT>::t V2;
    V2 *o = reinterpret_cast<V2 *>(coefM) + row * NSPLIT * 64 + l;
    const CP q = cell_prim(U_, J, c, p.K, e);
    const double dd[3] = {p.dxi, p.deta, p.dzeta};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const SC sp = split(q, met.m[3 * d][c], met.m[3 * d + 1][c], met.m[3 * d + 2][c], p.K, d != 0);
        const double h = dd[d];
        V2 *b = o + 7 * d * 64;
        b[0 * 64] = Pair2<CT>::make(sp.p12 / h, sp.p13 / h);
        b[1 * 64] = Pair2<CT>::make(sp.p14 / h, sp.p21 / h);
        b[2 * 64] = Pair2<CT>::make(sp.p31 / h, sp.p41 / h);
        b[3 * 64] = Pair2<CT>::make(sp.p51 / h, sp.p52 / h);
        b[4 * 64] = Pair2<CT>::make(sp.p53 / h, sp.p54 / h);
        b[5 * 64] = Pair2<CT>::make(sp.p11 / h, sp.p22 / h);
        b[6 * 64] = Pair2<CT>::make(sp.m11 / h, sp.m22 / h);
    }
    double m[9];
    load9(met, c, m);
    const DG D = diag(q, m, p, deltaT);
    double2 *od = reinterpret_cast<double2 *>(dgnM) + row * 2 * 64 + l;
    od[0] = make_double2(D.d11, D.d22);
    od[64] = make_double2(D.d51, 0.0);
    double *orr = mrM + row * 5 * 64 + l;
#pragma unroll
    for (int v = 0; v < 5; v++) orr[v * 64] = MR.v[v][c];
}

// the march-ordered increments NR back into their field, and into NL
// when NL.v[0] is not null (a sweep pair ends with NL = NR,
// Dual_time_stepping.cpp), interior cells only
__global__ void __launch_bounds__(256) k_march_scatter(Geo g, March M, const double *__restrict__ nrM, F5 NL,
                                                       F5 NR)
{
    const bool nl = NL.v[0] != nullptr;
    const MGather q = mgather_at(g, M);
    if (!q.ok) return;
    const long c = ix(g, 3 + q.il, 2 + q.jo, 2 + q.ko);
    const long o = march_row(M, q.T, q.t) * 5 * 64 + q.l;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double x = nrM[o + v * 64];
        if (nl) NL.v[v][c] = x;
        NR.v[v][c] = x;
    }
}

// per-lane constants of a march lane
struct MLane {
    int tid, jl, kl;
    bool col, edge_out, need_b, need_c;   // need_*: neighbour across a tile edge (exchange)
    int ib, ic;                           // LDS slots of the in-tile j and k neighbours
    unsigned oA, oB, oC;                  // record offsets (bytes) of the three upstream neighbours at t = 0
    long x0;                              // exchange-buffer index of (i = 3 - jl - kl, j, k), times 5
};

template <typename CT>
struct MArgs {
    Geo g;
    int S;
    __amdgpu_buffer_rsrc_t creg;   // coefficient rows from the wave's region base
    __amdgpu_buffer_rsrc_t drow;   // diagonal rows from (T, t = 0)
    __amdgpu_buffer_rsrc_t rrow;   // right-hand-side rows from (T, t = 0)
    char *xrow0;                   // increment row of (T, t = 0)
    double *xbuf;
    double *trash;                 // this wave's scratch line: exchange stores of lanes with nothing to publish
    int max_spin;
    long sxs5, dnb5, dnc5;   // exchange-buffer strides (times 5): one i-plane, the j and k neighbours
    int d;                   // -1 forward, +1 backward: towards the upstream neighbours
};

template <typename CT> struct MRowB { static constexpr long v = (long)NSPLIT * 64 * 2 * sizeof(CT); };
constexpr long DROWB = 2 * 64 * 16, FROWB = 5 * 64 * 8;

template <bool FWD, typename CT>
__device__ __forceinline__ int m_t(const MArgs<CT> &w, int s) { return FWD ? s : w.S - 1 - s; }

// buffer loads: a wave-uniform descriptor and scalar offset, a per-lane
// offset
template <typename CT> struct BufPair;
template <> struct BufPair<double> {
    static __device__ __forceinline__ double2 get(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
    {
        return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
    }
};
template <> struct BufPair<float> {
    static __device__ __forceinline__ float2 get(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
    {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
};

__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}

// one direction's 12 values (pairs 0..4 and the sweep's p- or m-pair) of
// the record at per-lane offset `vo`, row offset `so`, direction DIR
template <bool FWD, int DIR, typename CT>
__device__ __forceinline__ void load_dir_m(Dir12<CT> &s, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    constexpr unsigned P = 2 * sizeof(CT) * 64;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int p = 7 * DIR + (q < 5 ? q : (FWD ? 5 : 6));
        const typename Pair2<CT>::t t = BufPair<CT>::get(r, vo, so + p * P);
        s.v[2 * q] = t.x;
        s.v[2 * q + 1] = t.y;
    }
}

// the loads of step s (the last step's again for s = S, the prefetch past
// the end): the three upstream records, the own diagonal and R
template <bool FWD, typename CT>
__device__ __forceinline__ void m_load(Pref<CT> &P, const MArgs<CT> &w, const MLane &L, int s)
{
    const int t = m_t<FWD>(w, s < w.S ? s : w.S - 1);
    const unsigned so = (unsigned)(t * MRowB<CT>::v);   // wave-uniform
    // the i neighbour's record is the previous march slot of the lane;
    // at the first interior plane that neighbour is a ghost cell (zero
    // increment) and its slot holds no record: the own record stands in
    const int il = t - L.jl - L.kl;
    const bool inA = FWD ? il >= 1 : il <= w.g.nx - 2;
    const unsigned oa = inA ? L.oA : (unsigned)((long)L.oA - (long)w.d * MRowB<CT>::v);
    load_dir_m<FWD, 0>(P.A, w.creg, oa, so);
    load_dir_m<FWD, 1>(P.B, w.creg, L.oB, so);
    load_dir_m<FWD, 2>(P.C, w.creg, L.oC, so);
    const unsigned sd = (unsigned)(t * DROWB), sr = (unsigned)(t * FROWB);
    const double2 t0 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(w.drow, L.tid * 16, sd, 0));
    P.d11 = t0.x;
    P.d22 = t0.y;
    P.d51 = buf_f64(w.drow, L.tid * 16, sd + 1024);
#pragma unroll
    for (int v = 0; v < 5; v++) P.r[v] = buf_f64(w.rrow, L.tid * 8, sr + v * 512);
}

__device__ __forceinline__ void issue5m(const double *p, unsigned long long b[5])
{
#pragma unroll
    for (int v = 0; v < 5; v++) b[v] = aload(p + v);
}

// where `need`, re-poll the five entries of one neighbour until none holds
// the sentinel (bounded: `late` records an expired wait) and take them.
// The polled values are examined on every lane and path, so the wait for
// them happens here for all of them and a later poll into the same
// registers needs none.
__device__ __forceinline__ void settle5m(const double *p, bool need, unsigned long long b[5], double out[5],
                                         int max_spin, bool &late)
{
    bool pend = need & ((b[0] == SENTINEL) | (b[1] == SENTINEL) | (b[2] == SENTINEL) | (b[3] == SENTINEL) |
                        (b[4] == SENTINEL));
    if (pend) {
        for (int n = 0; pend && n < max_spin; n++) {
            __builtin_amdgcn_s_sleep(1);
            issue5m(p, b);
            pend = (b[0] == SENTINEL) | (b[1] == SENTINEL) | (b[2] == SENTINEL) | (b[3] == SENTINEL) |
                   (b[4] == SENTINEL);
        }
        late |= pend;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) out[v] = need ? __longlong_as_double((long long)b[v]) : out[v];
}

// one step: the records of step s are in register set Pc; step s+1's
// loads go into Pn
template <int TJ, int TK, bool FWD, int PAR, typename CT>
__device__ __forceinline__ void march_step(const MArgs<CT> &w, const MLane &L, int s, double prev[5],
                                           double (*sx)[5][TJ * TK], Pref<CT> &Pc, Pref<CT> &Pn,
                                           unsigned long long *nb, unsigned long long *nc, bool &late)
{
    const int t = m_t<FWD>(w, s);
    const int il = t - L.jl - L.kl;
    const bool act = L.col && il >= 0 && il < w.g.nx;
    __syncthreads();
    // exchange-buffer entry of this step's cell (only dereferenced where act)
    const long xc5 = L.x0 + (long)t * w.sxs5;
    double xb[5], xc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        xb[v] = sx[PAR ^ 1][v][L.ib];
        xc[v] = sx[PAR ^ 1][v][L.ic];
    }
    settle5m(w.xbuf + xc5 + w.dnb5, act && L.need_b, nb, xb, w.max_spin, late);
    settle5m(w.xbuf + xc5 + w.dnc5, act && L.need_c, nc, xc, w.max_spin, late);
    {   // prefetch for the next step: the tile-edge values (consuming lanes
        // with a cell there), then every lane's records
        const int tn = m_t<FWD>(w, s + 1);
        const int iln = tn - L.jl - L.kl;
        const bool actn = L.col && iln >= 0 && iln < w.g.nx;
        const long xn5 = L.x0 + (long)tn * w.sxs5;
        if (actn && L.need_b) issue5m(w.xbuf + xn5 + w.dnb5, nb);
        if (actn && L.need_c) issue5m(w.xbuf + xn5 + w.dnc5, nc);
        m_load<FWD>(Pn, w, L, s + 1);
    }
    // keep the prefetch ahead of the arithmetic that waits for this step's data
    __builtin_amdgcn_sched_barrier(0);
    const CT *A = Pc.A.v, *B = Pc.B.v, *Cc = Pc.C.v;
    CT xa[5], xbc[5], xcc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        xa[v] = (CT)prev[v];
        xbc[v] = (CT)xb[v];
        xcc[v] = (CT)xc[v];
    }
    double out[5];
    if (FWD) {
        const CT LL0 = (-A[10] * xa[0] - A[0] * xa[1] - A[1] * xa[2] - A[2] * xa[3]) +
                           (-B[10] * xbc[0] - B[0] * xbc[1] - B[1] * xbc[2] - B[2] * xbc[3]) +
                           (-Cc[10] * xcc[0] - Cc[0] * xcc[1] - Cc[1] * xcc[2] - Cc[2] * xcc[3]);
        const CT LL1 = (-A[3] * xa[0] - A[11] * xa[1]) + (-B[3] * xbc[0] - B[11] * xbc[1]) +
                           (-Cc[3] * xcc[0] - Cc[11] * xcc[1]);
        const CT LL2 = (-A[4] * xa[0] - A[11] * xa[2]) + (-B[4] * xbc[0] - B[11] * xbc[2]) +
                           (-Cc[4] * xcc[0] - Cc[11] * xcc[2]);
        const CT LL3 = (-A[5] * xa[0] - A[11] * xa[3]) + (-B[5] * xbc[0] - B[11] * xbc[3]) +
                           (-Cc[5] * xcc[0] - Cc[11] * xcc[3]);
        const CT LL4 = (-A[6] * xa[0] - A[7] * xa[1] - A[8] * xa[2] - A[9] * xa[3] - A[11] * xa[4]) +
                           (-B[6] * xbc[0] - B[7] * xbc[1] - B[8] * xbc[2] - B[9] * xbc[3] - B[11] * xbc[4]) +
                           (-Cc[6] * xcc[0] - Cc[7] * xcc[1] - Cc[8] * xcc[2] - Cc[9] * xcc[3] - Cc[11] * xcc[4]);
        out[0] = Pc.d11 * (Pc.r[0] - (double)LL0);
        out[1] = Pc.d22 * (Pc.r[1] - (double)LL1);
        out[2] = Pc.d22 * (Pc.r[2] - (double)LL2);
        out[3] = Pc.d22 * (Pc.r[3] - (double)LL3);
        out[4] = Pc.d51 * (Pc.r[0] - (double)LL0) + Pc.d22 * (Pc.r[4] - (double)LL4);
    } else {
        const CT UU0 = (+A[10] * xa[0] + A[0] * xa[1] + A[1] * xa[2] + A[2] * xa[3]) +
                           (+B[10] * xbc[0] + B[0] * xbc[1] + B[1] * xbc[2] + B[2] * xbc[3]) +
                           (+Cc[10] * xcc[0] + Cc[0] * xcc[1] + Cc[1] * xcc[2] + Cc[2] * xcc[3]);
        const CT UU1 = (+A[3] * xa[0] + A[11] * xa[1]) + (+B[3] * xbc[0] + B[11] * xbc[1]) +
                           (+Cc[3] * xcc[0] + Cc[11] * xcc[1]);
        const CT UU2 = (+A[4] * xa[0] + A[11] * xa[2]) + (+B[4] * xbc[0] + B[11] * xbc[2]) +
                           (+Cc[4] * xcc[0] + Cc[11] * xcc[2]);
        const CT UU3 = (+A[5] * xa[0] + A[11] * xa[3]) + (+B[5] * xbc[0] + B[11] * xbc[3]) +
                           (+Cc[5] * xcc[0] + Cc[11] * xcc[3]);
        const CT UU4 = (+A[6] * xa[0] + A[7] * xa[1] + A[8] * xa[2] + A[9] * xa[3] + A[11] * xa[4]) +
                           (+B[6] * xbc[0] + B[7] * xbc[1] + B[8] * xbc[2] + B[9] * xbc[3] + B[11] * xbc[4]) +
                           (+Cc[6] * xcc[0] + Cc[7] * xcc[1] + Cc[8] * xcc[2] + Cc[9] * xcc[3] + Cc[11] * xcc[4]);
        out[0] = Pc.r[0] - Pc.d11 * (double)UU0;
        out[1] = Pc.r[1] - Pc.d22 * (double)UU1;
        out[2] = Pc.r[2] - Pc.d22 * (double)UU2;
        out[3] = Pc.r[3] - Pc.d22 * (double)UU3;
        out[4] = Pc.r[4] - (Pc.d51 * (double)UU0 + Pc.d22 * (double)UU4);
    }
    // every lane stores into its own march slot; a slot with no cell is
    // never read as a value
    double *xo = reinterpret_cast<double *>(w.xrow0 + (long)t * FROWB) + L.tid;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        xo[v * 64] = out[v];
        // a column outside the slab holds the ghost value (zero) for its
        // neighbours: in the backward march it lies upstream of real ones
        sx[PAR][v][L.tid] = L.col ? out[v] : 0.0;
        prev[v] = act ? out[v] : prev[v];
    }
    // every lane stores: the edge lanes with a cell publish, the others
    // write the wave's scratch line, so no branch splits the stores and the
    // next step's wait for its records counts them exactly
    double *xe = act && L.edge_out ? w.xbuf + xc5 : w.trash;
#pragma unroll
    for (int v = 0; v < 5; v++)
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(xe + v),
                           (unsigned long long)__double_as_longlong(out[v]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// coefA/dgnA/RmA/XmA: the padded march arrays; Xg: the increment field,
// whose ghost plane upstream in i starts each column
template <int TJ, int TK, bool FWD, typename CT>
__global__ void __launch_bounds__(TJ *TK) k_lusgs_march(Geo g, March M, const CT *__restrict__ coefA,
                                                      const double *__restrict__ dgnA, const double *RmA,
                                                      double *XmA, C5 Xg, double *xbuf, double *trash,
                                                      const int *__restrict__ order, int ntiles, int *ticket,
                                                      int *err, int max_spin, int slack)
{
    static_assert(TJ * TK == 64, "one wave per tile");
    __shared__ int s_t;
    __shared__ double sx[2][5][TJ * TK];
    MLane L;
    L.tid = threadIdx.x;
    L.kl = L.tid % TK;
    L.jl = L.tid / TK;
    if (L.tid == 0) s_t = atomicAdd(ticket, 1);
    __syncthreads();
    // forward: anti-diagonals ascending; backward: the same list reversed
    const int tk = __builtin_amdgcn_readfirstlane(s_t);
    const int packed = order[FWD ? tk : ntiles - 1 - tk];
    const int TJi = packed >> 16, TKi = packed & 0xffff;
    const int T = TJi * M.ntk + TKi;
    const int jo = TJi * TJ + L.jl, ko = TKi * TK + L.kl;
    L.col = (jo < g.ny) && (ko < g.nz);
    const int joc = jo < g.ny ? jo : g.ny - 1, koc = ko < g.nz ? ko : g.nz - 1;
    const long c0 = ix(g, 0, 2 + joc, 2 + koc);
    const int d = FWD ? -1 : 1;
    // the neighbour records: tile, lane and march-index offset of the j and
    // k neighbours (the own record where there is none)
    int Tb, lb, tbo, Tc, lc, tco;
    if (FWD) {
        L.need_b = L.jl == 0;
        L.need_c = L.kl == 0;
        L.edge_out = (L.jl == TJ - 1) || (L.kl == TK - 1);
        // j-1: the lane above in this tile, or the last row of the tile above
        if (L.jl >= 1) { Tb = T; lb = L.tid - TK; tbo = -1; }
        else if (TJi >= 1) { Tb = T - M.ntk; lb = (TJ - 1) * TK + L.kl; tbo = TJ - 1; }
        else { Tb = T; lb = L.tid; tbo = 0; }
        if (L.kl >= 1) { Tc = T; lc = L.tid - 1; tco = -1; }
        else if (TKi >= 1) { Tc = T - 1; lc = L.jl * TK + TK - 1; tco = TK - 1; }
        else { Tc = T; lc = L.tid; tco = 0; }
        L.ib = L.jl >= 1 ? L.tid - TK : L.tid;
        L.ic = L.kl >= 1 ? L.tid - 1 : L.tid;
    } else {
        L.need_b = L.jl == TJ - 1;
        L.need_c = L.kl == TK - 1;
        L.edge_out = (L.jl == 0) || (L.kl == 0);
        // j+1: the lane below in this tile, or the first row of the tile below
        if (L.jl <= TJ - 2 && jo + 1 < g.ny) { Tb = T; lb = L.tid + TK; tbo = 1; }
        else if (L.jl == TJ - 1 && TJi + 1 < M.ntj) { Tb = T + M.ntk; lb = L.kl; tbo = -(TJ - 1); }
        else { Tb = T; lb = L.tid; tbo = 0; }
        if (L.kl <= TK - 2 && ko + 1 < g.nz) { Tc = T; lc = L.tid + 1; tco = 1; }
        else if (L.kl == TK - 1 && TKi + 1 < M.ntk) { Tc = T + 1; lc = L.jl * TK; tco = -(TK - 1); }
        else { Tc = T; lc = L.tid; tco = 0; }
        L.ib = L.jl <= TJ - 2 ? L.tid + TK : L.tid;
        L.ic = L.kl <= TK - 2 ? L.tid + 1 : L.tid;
    }
    // records: a wave-uniform region base (the lowest row the sweep reads
    // from) and per-lane byte offsets of the rows at t = 0
    typedef typename Pair2<CT>::t V2;
    constexpr long RB = MRowB<CT>::v;
    const int Tlo = FWD ? (T >= M.ntk ? T - M.ntk : 0) : T;
    const long rlo = (long)Tlo * M.S;   // padded row of (Tlo, t = -MPAD)
    MArgs<CT> w;
    w.g = g;
    w.S = M.S;
    constexpr int F = 0x00020000;   // raw buffer, 32-bit data format
    w.creg = __builtin_amdgcn_make_buffer_rsrc((void *)(reinterpret_cast<const char *>(coefA) + rlo * RB), (short)0, -1, F);
    L.oA = (unsigned)((march_row(M, T, d) - rlo) * RB + L.tid * (long)sizeof(V2));
    L.oB = (unsigned)((march_row(M, Tb, tbo) - rlo) * RB + lb * (long)sizeof(V2));
    L.oC = (unsigned)((march_row(M, Tc, tco) - rlo) * RB + lc * (long)sizeof(V2));
    w.drow = __builtin_amdgcn_make_buffer_rsrc((void *)(reinterpret_cast<const char *>(dgnA) + march_row(M, T, 0) * DROWB),
                                               (short)0, -1, F);
    w.rrow = __builtin_amdgcn_make_buffer_rsrc((void *)(reinterpret_cast<const char *>(RmA) + march_row(M, T, 0) * FROWB),
                                               (short)0, -1, F);
    w.xrow0 = reinterpret_cast<char *>(XmA) + march_row(M, T, 0) * FROWB;
    w.xbuf = xbuf;
    w.trash = trash + (long)blockIdx.x * 256;
    w.max_spin = max_spin;
    const long sxs = (long)g.Ym * g.Zm;
    w.sxs5 = 5 * sxs;
    w.dnb5 = 5L * d * g.Zm;
    w.dnc5 = 5L * d;
    w.d = d;
    L.x0 = 5 * (c0 + (long)(3 - L.jl - L.kl) * sxs);
    const int S = M.S;
    double prev[5];
    bool late = false;
    Pref<CT> P0, P1;
    unsigned long long nb[5], nc[5];
    // Start `slack` march steps later than the dependencies require: wait
    // until the upstream tiles have published the exchange entries of step
    // `slack`.  Tiles march at the same rate, so the lead persists and each
    // step's poll then finds its entries written instead of re-polling.
    if (slack > 0 && slack < S) {
        const int ts = m_t<FWD>(w, slack);
        const int ils = ts - L.jl - L.kl;
        const bool acts = L.col && ils >= 0 && ils < g.nx;
        const long xs5 = L.x0 + (long)ts * w.sxs5;
        unsigned long long sb[5] = {0, 0, 0, 0, 0}, sc[5] = {0, 0, 0, 0, 0};
        double tmp[5] = {0, 0, 0, 0, 0};
        if (acts && L.need_b) issue5m(xbuf + xs5 + w.dnb5, sb);
        if (acts && L.need_c) issue5m(xbuf + xs5 + w.dnc5, sc);
        settle5m(xbuf + xs5 + w.dnb5, acts && L.need_b, sb, tmp, max_spin, late);
        settle5m(xbuf + xs5 + w.dnc5, acts && L.need_c, sc, tmp, max_spin, late);
    }
    {
        // the ghost plane upstream in i (zero increments)
        const long cg = c0 + (long)(FWD ? 2 : g.gend + 1) * sxs;
#pragma unroll
        for (int v = 0; v < 5; v++) prev[v] = Xg.v[v][cg];
        const int t0 = m_t<FWD>(w, 0);
        const int il0 = t0 - L.jl - L.kl;
        const bool act0 = L.col && il0 >= 0 && il0 < g.nx;
        const long x5 = L.x0 + (long)t0 * w.sxs5;
        if (act0 && L.need_b) issue5m(xbuf + x5 + w.dnb5, nb);
        if (act0 && L.need_c) issue5m(xbuf + x5 + w.dnc5, nc);
        m_load<FWD>(P0, w, L, 0);
    }
    for (int s = 0; s < S; s += 2) {
        march_step<TJ, TK, FWD, 0>(w, L, s, prev, sx, P0, P1, nb, nc, late);
        if (s + 1 < S) march_step<TJ, TK, FWD, 1>(w, L, s + 1, prev, sx, P1, P0, nb, nc, late);
    }
    if (late) atomicOr(err, 1);
}

// ==================================================================
// Schedule 4: the march without coefficient records.
//
// The sweep reads, per cell and step, the cell's state U_, its Jacobian and
// 9 metrics and its right-hand side (MR forward, NL backward), all held in
// march order (J and the metrics gathered once, they do not change), and
// computes the cell's split coefficients and diagonal itself, one step
// before the step that needs them, so that this arithmetic runs while the
// step's exchange entries are on their way.  What travels between cells
// are the products of an increment with the split block its downstream
// neighbour needs along each direction -- the per-direction sums of
// Dual_time_stepping.cpp:1172-1232 (forward: plus blocks, negated) and
// 1529-1590 (backward: minus diagonal entries) -- already divided by the
// grid spacing: along i in a register of the same lane, along j and k
// through LDS inside a tile and through two exchange buffers (j products,
// k products) across tile edges.  Compared with the plane order (which
// follows the reference's expressions literally) the divisions by U1, rho,
// the grid spacings and the diagonal's denominators are multiplications by
// reciprocals: results agree to rounding.
//
// March-ordered arrays hold 16-byte pairs [row][pair][lane]: the geometry
// (J, m0) (m1, m2) (m3, m4) (m5, m6) (m7, m8), and five-value fields (U_,
// MR, NL, NR) as (v0, v1) (v2, v3) (v4, -).
constexpr int G4P = 5, F4P = 3;
// Extra march steps a tile waits, before it starts, for its upstream tiles
// to run ahead.  With the record exchange no wait is fastest (256x192x256
// forward sweep 1.77 ms at 0, 1.82 at 1, 1.90 at 2, 1.98 at 3, 2.27 at 6):
// a tile starts at once and its first polls spin until the records arrive.
#ifndef CGLI_FUSED_SLACK
#define CGLI_FUSED_SLACK 0
#endif
constexpr int FUSED_SLACK = CGLI_FUSED_SLACK;
constexpr long G4ROWB = (long)G4P * 64 * 16, F4ROWB = (long)F4P * 64 * 16;

long march4_geo_pairs(const Geo &g) { return march_rows(g) * G4P * 64; }
long march4_field_pairs(const Geo &g) { return march_rows(g) * F4P * 64; }

// U_ (when Um), a right-hand side R (when Rm) and the geometry (when Gm)
// into march order; the slots of cells only (the others keep their fill)
__global__ void __launch_bounds__(256) k_gather4(Geo g, March M, C5 U_, C5 R, const double *__restrict__ J, M9 met,
                                                 double2 *__restrict__ Um, double2 *__restrict__ Rm,
                                                 double2 *__restrict__ Gm)
{
    const MGather q = mgather_at(g, M);
    if (!q.ok) return;
    const long c = ix(g, 3 + q.il, 2 + q.jo, 2 + q.ko);
    const long row = march_row(M, q.T, q.t);
    const long o = row * F4P * 64 + q.l;
    if (Um) {
        Um[o] = make_double2(U_.v[0][c], U_.v[1][c]);
        Um[o + 64] = make_double2(U_.v[2][c], U_.v[3][c]);
        Um[o + 128] = make_double2(U_.v[4][c], 0.0);
    }
    if (Rm) {
        Rm[o] = make_double2(R.v[0][c], R.v[1][c]);
        Rm[o + 64] = make_double2(R.v[2][c], R.v[3][c]);
        Rm[o + 128] = make_double2(R.v[4][c], 0.0);
    }
    if (Gm) {
        const long og = row * G4P * 64 + q.l;
        Gm[og] = make_double2(J[c], met.m[0][c]);
        Gm[og + 64] = make_double2(met.m[1][c], met.m[2][c]);
        Gm[og + 128] = make_double2(met.m[3][c], met.m[4][c]);
        Gm[og + 192] = make_double2(met.m[5][c], met.m[6][c]);
        Gm[og + 256] = make_double2(met.m[7][c], met.m[8][c]);
    }
}

// Schedule 4's update and norm (k_update_norm's arithmetic) in the march
// mapping of k_gather4 / k_scatter4: the increments NR and U_ are read in
// march order, where the sweep left them and the gather put them this
// iteration, so NR goes to no field (its field is written from march order
// where read); the new U_ is written to its field and, in place, to its
// march-ordered copy, which the next inner iteration's gather then need not
// rebuild.  Each block writes 5 partial sums of its positions in a fixed
// order.
__global__ void __launch_bounds__(256) k_update4(Geo g, March M, Phys p, F5 U_, const double *__restrict__ J,
                                                 double2 *__restrict__ Um, const double2 *__restrict__ Xm,
                                                 F5 Res, int write_res, double *partials)
{
    const MGather q = mgather_at(g, M);
    double x[5] = {0, 0, 0, 0, 0};
    if (q.ok) {
        const long c = ix(g, 3 + q.il, 2 + q.jo, 2 + q.ko);
        const long o = march_row(M, q.T, q.t) * F4P * 64 + q.l;
        const double2 u01 = Um[o], u23 = Um[o + 64], u4 = Um[o + 128];
        const double2 n01 = Xm[o], n23 = Xm[o + 64], n4 = Xm[o + 128];
        const double K = p.K;
        const double Jc = J[c];
        // quotients that share a divisor (U1, Jc) through its reciprocal
        // and div_r: the division's bits
        const double U1 = u01.x, iU1 = 1.0 / U1, iJc = 1.0 / Jc;
        const double rhoold = U1 * Jc;
        const double Uold = div_r(u01.y, U1, iU1);
        const double Vold = div_r(u23.x, U1, iU1);
        const double Wold = div_r(u23.y, U1, iU1);
        const double VVold = Uold * Uold + Vold * Vold + Wold * Wold;
        const double Pold = (u4.x * Jc - 0.5 * rhoold * VVold) * (K - 1);
        const double Told = Pold / rhoold;
        const double P = Pold + n01.x * Jc;
        const double Uu = Uold + n01.y * Jc;
        const double V = Vold + n23.x * Jc;
        const double W = Wold + n23.y * Jc;
        const double T = Told + n4.x * Jc;
        const double rho = P / T;
        const double w0 = div_r(rho, Jc, iJc), w1 = div_r(rho * Uu, Jc, iJc), w2 = div_r(rho * V, Jc, iJc);
        const double w3 = div_r(rho * W, Jc, iJc);
        const double w4 = div_r(P / (K - 1) + 0.5 * rho * (Uu * Uu + V * V + W * W), Jc, iJc);
        U_.v[0][c] = w0;
        U_.v[1][c] = w1;
        U_.v[2][c] = w2;
        U_.v[3][c] = w3;
        U_.v[4][c] = w4;
        Um[o] = make_double2(w0, w1);
        Um[o + 64] = make_double2(w2, w3);
        Um[o + 128] = make_double2(w4, 0.0);
        x[0] = (rho - rhoold) * (rho - rhoold);
        x[1] = (Uu - Uold) * (Uu - Uold);
        x[2] = (V - Vold) * (V - Vold);
        x[3] = (W - Wold) * (W - Wold);
        x[4] = (T - Told) * (T - Told);
        if (write_res) {
#pragma unroll
            for (int v = 0; v < 5; v++) Res.v[v][c] = x[v];
        }
    }
    double o5[5];
    block_sum<5>(x, o5);
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 5; v++) partials[5 * (long)blockIdx.x + v] = o5[v];
}

long update4_blocks(const Geo &g) { return mgather_grid(march_of(g)).x; }

int launch_update4(const Geo &g, const Phys &p, F5 U_, const double *J, void *Um, const void *Xm, F5 Res,
                   int write_res, double *partials, hipStream_t s)
{
    const March M = march_of(g);
    const dim3 gr = mgather_grid(M);
    hipLaunchKernelGGL(k_update4, gr, dim3(256), 0, s, g, M, p, U_, J, static_cast<double2 *>(Um),
                       static_cast<const double2 *>(Xm), Res, write_res, partials);
    return (int)gr.x;
}

// march-ordered increments back into NR (and NL when NL.v[0] is not null)
__global__ void __launch_bounds__(256) k_scatter4(Geo g, March M, const double2 *__restrict__ Xm, F5 NL, F5 NR)
{
    const bool nl = NL.v[0] != nullptr;
    const MGather q = mgather_at(g, M);
    if (!q.ok) return;
    const long c = ix(g, 3 + q.il, 2 + q.jo, 2 + q.ko);
    const long o = march_row(M, q.T, q.t) * F4P * 64 + q.l;
    const double2 a = Xm[o], b = Xm[o + 64], d = Xm[o + 128];
    const double x[5] = {a.x, a.y, b.x, b.y, d.x};
#pragma unroll
    for (int v = 0; v < 5; v++) {
        if (nl) NL.v[v][c] = x[v];
        NR.v[v][c] = x[v];
    }
}

// the loads of one step and the coefficients computed from them
struct Ld4 { double2 g[G4P], u[F4P], r[F4P]; };
// c: the compact coefficient set of the cell (coef_compact), r its right-hand
// side, m its nine metrics
struct Cf4 { double c[21]; double r[5]; double m[9]; };

struct F4Lane {
    int tid, jl, kl;
    bool col;                 // the lane's (j, k) column lies in the slab
    bool hasb, hasc;          // in-tile upstream neighbour along j / k (in the slab)
    bool need_b, need_c;      // upstream neighbour across a tile edge (exchange)
    bool out_b, out_c;        // downstream neighbour across a tile edge (publish)
    int ib, ic;               // LDS slots of the in-tile upstream neighbours (else the zero entry TJ * TK)
    unsigned pb, pc;          // byte offsets of the polled j / k records at step 0 (XR_OOB: none)
    unsigned ob, oc;          // byte offsets of the published j / k records at step 0
};

struct M4Args {
    Geo g;
    int S;
    double K, iK, e, deltaT;
    double h[3];              // 1 / dxi, 1 / deta, 1 / dzeta
    __amdgpu_buffer_rsrc_t grow, urow, rrow;   // rows from (T, t = 0)
    char *xrow0;              // output row of (T, t = 0)
    __amdgpu_buffer_rsrc_t rb, rc;   // the j and k exchange records
    int dtb, dtc;             // a polled record's step minus the polling cell's step
    int max_spin;
    double2 gcol[G4P];        // GC: the column's geometry, the same in every plane
};

// Tile-edge exchange records.  A tile's cell on its downstream j edge
// publishes, at its march step t, the product of its increment with its j
// block (five doubles and a zero pad, 48 bytes) as entry e = its k position of
// record (T, t) in the j buffer; the k edge likewise into the k buffer.  The
// eight entries of one record are contiguous, so an edge's publication and
// its consumer's poll touch three 128-byte lines.  The consumer cell of the
// same column one tile downstream needs that entry at its own step t - dtb
// (a tile edge's cells are tile-width steps apart in the skewed march).
// All accesses are 16-byte buffer operations at device scope (write-through
// stores, L2-bypassing loads); lanes with nothing to publish or poll use an
// offset past the buffer, which the hardware discards (loads return zero),
// so every lane issues the same instructions and no memory request is made
// for them.  Each 8-byte value is validated on its own against the sentinel
// the buffers are filled with before each sweep.
typedef unsigned int xr_u32x4 __attribute__((ext_vector_type(4)));
constexpr unsigned XR_OOB = 0x80000000u;
constexpr int XR_CPOL = 16;                           // sc1: device scope
constexpr unsigned XR_ENTRY = 48, XR_STEP = 8 * XR_ENTRY;

__device__ __forceinline__ unsigned long long xr_lo(xr_u32x4 q) { return q.x | ((unsigned long long)q.y << 32); }
__device__ __forceinline__ unsigned long long xr_hi(xr_u32x4 q) { return q.z | ((unsigned long long)q.w << 32); }

// the five values of an entry: two 16-byte loads and one 8-byte load (not
// the pad: a loaded register left unused may be reused by the compiler while
// the load is in flight, which costs a wait for every outstanding load)
__device__ __forceinline__ void rec_issue(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned long long b[6])
{
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const xr_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16u * p, 0, XR_CPOL);
        b[2 * p] = xr_lo(q);
        b[2 * p + 1] = xr_hi(q);
    }
    b[4] = __builtin_bit_cast(unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(r, vo + 32u, 0, XR_CPOL));
    b[5] = 0;
}

// where `need`, re-poll the record entry until none of its five values holds
// the sentinel (bounded: `late` records an expired wait) and take them
__device__ __forceinline__ void rec_settle(__amdgpu_buffer_rsrc_t r, unsigned vo, bool need, unsigned long long b[6],
                                           double out[5], int max_spin, bool &late)
{
    bool pend = need & ((b[0] == SENTINEL) | (b[1] == SENTINEL) | (b[2] == SENTINEL) | (b[3] == SENTINEL) |
                        (b[4] == SENTINEL));
    if (pend) {
        for (int n = 0; pend && n < max_spin; n++) {
            __builtin_amdgcn_s_sleep(1);
            rec_issue(r, vo, b);
            pend = (b[0] == SENTINEL) | (b[1] == SENTINEL) | (b[2] == SENTINEL) | (b[3] == SENTINEL) |
                   (b[4] == SENTINEL);
        }
        late |= pend;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) out[v] = need ? __longlong_as_double((long long)b[v]) : out[v];
}

__device__ __forceinline__ void rec_store(__amdgpu_buffer_rsrc_t r, unsigned vo, const double p[5])
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(xr_u32x4, make_double2(p[0], p[1])), r, vo, 0, XR_CPOL);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(xr_u32x4, make_double2(p[2], p[3])), r, vo + 16u, 0,
                                           XR_CPOL);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(xr_u32x4, make_double2(p[4], 0.0)), r, vo + 32u, 0,
                                           XR_CPOL);
}

// the record offsets of step t for a lane: polled (where act and needed) and
// published (where act and on the edge)
__device__ __forceinline__ unsigned rec_poll(unsigned base, bool on, int t, int dt)
{
    return on ? base + (unsigned)(t + dt) * XR_STEP : XR_OOB;
}
__device__ __forceinline__ unsigned rec_pub(unsigned base, bool on, int t)
{
    return on ? base + (unsigned)t * XR_STEP : XR_OOB;
}

__device__ __forceinline__ double2 ld_pair(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
}

template <bool FWD>
__device__ __forceinline__ int m4_t(const M4Args &w, int s)
{
    const int sc = s < w.S ? s : w.S - 1;   // past the end: the last step's row again
    return FWD ? sc : w.S - 1 - sc;
}

// GC: the grid's metrics do not vary along i (the channel's mapping does not
// depend on xi), so each lane holds its column's geometry in registers for
// the whole sweep instead of loading it at every step
template <bool FWD, bool GC>
__device__ __forceinline__ void load4(Ld4 &L, const M4Args &w, unsigned vo, int s)
{
    const int t = m4_t<FWD>(w, s);
    const unsigned sg = (unsigned)(t * G4ROWB), sf = (unsigned)(t * F4ROWB);
#pragma unroll
    for (int p = 0; p < G4P; p++) L.g[p] = GC ? w.gcol[p] : ld_pair(w.grow, vo, sg + p * 1024);
#pragma unroll
    for (int p = 0; p < F4P; p++) L.u[p] = ld_pair(w.urow, vo, sf + p * 1024);
#pragma unroll
    for (int p = 0; p < F4P; p++) L.r[p] = ld_pair(w.rrow, vo, sf + p * 1024);
}

// primitives (Dual_time_stepping.cpp:759-769), the split coefficients of the
// three directions (1054-1170, 1384-1524) and the diagonal (856-895) of one
// cell; the directions' spectral radii are shared by the split and the
// diagonal.  The compact set holds per direction the entries that do not
// multiply a metric (a0 = g11 or m11, a5 = g22 or m22, a8) and the three
// factors that do (pk, ir, bt), then d11, d22, d51; prod_c applies a
// direction's block to an increment from it.  The division by K is a
// product with 1/K and the two diagonal reciprocals come from one division,
// so the set agrees with the reference's to rounding.
constexpr int CF_N = 21;

template <bool FWD>
__device__ __forceinline__ void coef_compact(double c[CF_N], double J, const double m[9], const double U[5],
                                             const M4Args &w)
{
    const double K = w.K;
    const double iU1 = 1.0 / U[0];
    const double rho = U[0] * J;
    const double u = U[1] * iU1, v = U[2] * iU1, ww = U[3] * iU1;
    const double VV = u * u + v * v + ww * ww;
    const double P = (U[4] * J - 0.5 * rho * VV) * (K - 1);
    const double irho = 1.0 / rho;
    const double C = K * P * irho;
    const double T = P * irho;
    const double rr = VV / C;
    const double b = rr > w.e ? rr : w.e;
    double lam = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double mx = m[3 * d], my = m[3 * d + 1], mz = m[3 * d + 2];
        const double Un = mx * u + my * v + mz * ww;
        const double S = sqrt(Un * Un * (b - 1) * (b - 1) + 4 * b * C * (mx * mx + my * my + mz * mz));
        const double l = ((b + 1) * fabs(Un) + S) * 0.5;
        const double h = w.h[d];
        lam += l * h;
        const double dd = fabs(l);
        double *a = c + 6 * d;
        a[0] = FWD ? 0.5 * (b * Un + dd) * h : 0.5 * (b * Un - dd) * h;
        a[1] = FWD ? 0.5 * (Un + dd) * h : 0.5 * (Un - dd) * h;
        a[2] = 0.5 * P * K * b * h;
        const double ir = 0.5 * irho * h;
        a[3] = ir;
        a[4] = Un * (K - 1) * (b - 1) * (ir * w.iK);
        a[5] = 0.5 * b * T * (K - 1) * h;
    }
    // the two diagonal reciprocals from one division
    const double dt = w.deltaT;
    const double dA = 3 + 2 * lam * dt, dB = 3 * b + 2 * lam * dt;
    const double rAB = 1.0 / (dA * dB);
    const double rA = dB * rAB, rB = dA * rAB;
    c[18] = 2 * dt * rB;
    c[19] = 2 * dt * rA;
    c[20] = -(6 * (K - 1) / K) * (b - 1) * dt * rA * rB * irho;
}

template <bool FWD>
__device__ __forceinline__ void coef4(Cf4 &C, const Ld4 &L, const M4Args &w)
{
    const double m[9] = {L.g[0].y, L.g[1].x, L.g[1].y, L.g[2].x, L.g[2].y, L.g[3].x, L.g[3].y, L.g[4].x, L.g[4].y};
    const double U[5] = {L.u[0].x, L.u[0].y, L.u[1].x, L.u[1].y, L.u[2].x};
    coef_compact<FWD>(C.c, L.g[0].x, m, U, w);
#pragma unroll
    for (int q = 0; q < 9; q++) C.m[q] = m[q];
    C.r[0] = L.r[0].x;
    C.r[1] = L.r[0].y;
    C.r[2] = L.r[1].x;
    C.r[3] = L.r[1].y;
    C.r[4] = L.r[2].x;
}

// the product of increment x with one direction's block, from that
// direction's compact coefficients q (six, from c + 6 d) and metrics m3
// (three, from m + 3 d): forward the negated plus block (a lower
// neighbour's contribution), backward the block with the minus diagonal
// entries (an upper neighbour's).  The block's entries are q0 and q1 on the
// diagonal, q2 m, q3 m, q5 m in the first row, first column and last row,
// q4 in the corner (Dual_time_stepping.cpp:1054-1170, 1384-1524); the
// metric entries enter through m . (x1, x2, x3) and q3 x0, with fused
// multiply-adds, so the product agrees with the entry-by-entry one to
// rounding.
template <bool FWD>
__device__ __forceinline__ void prod_c(const double *q, const double *m3, const double x[5], double y[5])
{
    // the forward negation as operand signs (free source modifiers)
    const double sg = FWD ? -1.0 : 1.0;
    const double md = fma(m3[2], x[3], fma(m3[1], x[2], m3[0] * x[1]));
    const double qx = (sg * q[3]) * x[0];
    const double q1 = sg * q[1];
    y[0] = fma(sg * q[2], md, (sg * q[0]) * x[0]);
    y[1] = fma(qx, m3[0], q1 * x[1]);
    y[2] = fma(qx, m3[1], q1 * x[2]);
    y[3] = fma(qx, m3[2], q1 * x[3]);
    y[4] = fma(q1, x[4], fma(sg * q[5], md, (sg * q[4]) * x[0]));
}

// both directions' polls of a step settled: until neither record entry
// holds the sentinel (bounded: `late` records an expired wait), then their
// values added to the in-tile products.  A lane that polls nothing loaded
// from past the buffer and holds zeros (never the sentinel); a lane that
// polls has its in-tile slot at the zero entry, so the sums are the one
// value each lane needs.
__device__ __forceinline__ bool rec_pending(const unsigned long long b[6])
{
    return (b[0] == SENTINEL) | (b[1] == SENTINEL) | (b[2] == SENTINEL) | (b[3] == SENTINEL) | (b[4] == SENTINEL);
}
__device__ __forceinline__ void rec_settle2(const M4Args &w, unsigned vb, unsigned vc, unsigned long long nb[6],
                                            unsigned long long nc[6], double yb[5], double yc[5], bool &late)
{
    bool pend = rec_pending(nb) | rec_pending(nc);
    if (pend) {
        for (int n = 0; pend && n < w.max_spin; n++) {
            __builtin_amdgcn_s_sleep(1);
            rec_issue(w.rb, vb, nb);
            rec_issue(w.rc, vc, nc);
            pend = rec_pending(nb) | rec_pending(nc);
        }
        late |= pend;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
        yb[v] += __longlong_as_double((long long)nb[v]);
        yc[v] += __longlong_as_double((long long)nc[v]);
    }
}

// one march step s; C holds step s's coefficients, Ln the loads of step
// s + 1 (issued a step ago), Lf receives the loads of step s + 2; C leaves
// with step s + 1's coefficients.  nb / nc hold the exchange polls of step
// s (issued two steps ago) and receive those of step s + 2.  The in-tile
// products come from LDS (a lane with no in-tile upstream neighbour reads
// the zero entry TJ * TK); the compiler interleaves the next coefficients
// with this step's dependent chain.
template <int TJ, int TK, bool FWD, bool GC, int PAR>
__device__ __forceinline__ void step4(const M4Args &w, const F4Lane &L, int s, Cf4 &C, const Ld4 &Ln, Ld4 &Lf,
                                      double sa[5], double (*sB)[5][TJ * TK + 1], double (*sC)[5][TJ * TK + 1],
                                      unsigned long long nb[6], unsigned long long nc[6], bool &late)
{
    const int t = FWD ? s : w.S - 1 - s;
    const int il = t - L.jl - L.kl;
    const bool act = L.col && il >= 0 && il < w.g.nx;
    __syncthreads();
    double yb[5], yc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        yb[v] = sB[PAR ^ 1][v][L.ib];
        yc[v] = sC[PAR ^ 1][v][L.ic];
    }
    rec_settle2(w, rec_poll(L.pb, act && L.need_b, t, w.dtb), rec_poll(L.pc, act && L.need_c, t, w.dtc), nb, nc,
                yb, yc, late);
    const double *c = C.c;
    const double d11 = c[18], d22 = c[19], d51 = c[20];
    double x[5];
    if (FWD) {
        const double LL0 = sa[0] + yb[0] + yc[0], LL1 = sa[1] + yb[1] + yc[1], LL2 = sa[2] + yb[2] + yc[2];
        const double LL3 = sa[3] + yb[3] + yc[3], LL4 = sa[4] + yb[4] + yc[4];
        x[0] = d11 * (C.r[0] - LL0);
        x[1] = d22 * (C.r[1] - LL1);
        x[2] = d22 * (C.r[2] - LL2);
        x[3] = d22 * (C.r[3] - LL3);
        x[4] = d51 * (C.r[0] - LL0) + d22 * (C.r[4] - LL4);
    } else {
        const double UU0 = sa[0] + yb[0] + yc[0], UU1 = sa[1] + yb[1] + yc[1], UU2 = sa[2] + yb[2] + yc[2];
        const double UU3 = sa[3] + yb[3] + yc[3], UU4 = sa[4] + yb[4] + yc[4];
        x[0] = C.r[0] - d11 * UU0;
        x[1] = C.r[1] - d22 * UU1;
        x[2] = C.r[2] - d22 * UU2;
        x[3] = C.r[3] - d22 * UU3;
        x[4] = C.r[4] - (d51 * UU0 + d22 * UU4);
    }
    const double *m = C.m;
    double pa[5], pb[5], pc[5];
    prod_c<FWD>(c, m, x, pa);
    prod_c<FWD>(c + 6, m + 3, x, pb);
    prod_c<FWD>(c + 12, m + 6, x, pc);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        // read only by in-tile neighbours whose column lies in the slab
        sB[PAR][v][L.tid] = pb[v];
        sC[PAR][v][L.tid] = pc[v];
        sa[v] = act ? pa[v] : sa[v];
    }
    {   // the exchange polls and the loads two steps ahead
        const int tn = FWD ? s + 2 : w.S - 3 - s;
        const int iln = tn - L.jl - L.kl;
        const bool actn = s + 2 < w.S && L.col && iln >= 0 && iln < w.g.nx;
        rec_issue(w.rb, rec_poll(L.pb, actn && L.need_b, tn, w.dtb), nb);
        rec_issue(w.rc, rec_poll(L.pc, actn && L.need_c, tn, w.dtc), nc);
        load4<FWD, GC>(Lf, w, (unsigned)L.tid * 16u, s + 2);
    }
    // loads before stores: the vector-memory counter completes in issue
    // order, so a wait for these loads must not also wait for the stores
    __builtin_amdgcn_sched_barrier(0);
    // every lane stores its slot (a slot with no cell is never read as a
    // value) and issues the six record stores, so no branch splits them
    double2 *xo = reinterpret_cast<double2 *>(w.xrow0 + (long)t * F4ROWB) + L.tid;
    xo[0] = make_double2(x[0], x[1]);
    xo[64] = make_double2(x[2], x[3]);
    xo[128] = make_double2(x[4], 0.0);
    rec_store(w.rb, rec_pub(L.ob, act && L.out_b, t), pb);
    rec_store(w.rc, rec_pub(L.oc, act && L.out_c, t), pc);
    // the next step's coefficients while all of it is in flight
    coef4<FWD>(C, Ln, w);
}

// the x-direction product of the ghost cell upstream in i (its increment Xg
// from the field: zero on a periodic or end slab), which starts each column
template <bool FWD>
__device__ __forceinline__ void ghost_product4(const M4Args &w, C5 U_, const double *__restrict__ J, M9 met, C5 Xg,
                                               long cg, double y[5])
{
    const double m[9] = {met.m[0][cg], met.m[1][cg], met.m[2][cg], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double U[5] = {U_.v[0][cg], U_.v[1][cg], U_.v[2][cg], U_.v[3][cg], U_.v[4][cg]};
    double c[CF_N];
    coef_compact<FWD>(c, J[cg], m, U, w);
    const double x[5] = {Xg.v[0][cg], Xg.v[1][cg], Xg.v[2][cg], Xg.v[3][cg], Xg.v[4][cg]};
    prod_c<FWD>(c, m, x, y);
}

// the lane's place in its tile and the tile's march arguments (both
// schedule-4 kernels); returns the tile's index T in the tile grid and the
// field index of the lane's column at i = 0 in c0
template <int TJ, int TK, bool FWD>
__device__ __forceinline__ int fused_setup(F4Lane &L, M4Args &w, long &c0, int lane, int packed, const Geo &g,
                                           const March &M, double K, double e, double deltaT, double hx, double hy,
                                           double hz, const double2 *Gm, const double2 *Um, const double2 *Rm,
                                           double2 *Xm, double2 *rec, unsigned rec_bytes, int max_spin)
{
    L.tid = lane;
    L.kl = L.tid % TK;
    L.jl = L.tid / TK;
    const int TJi = packed >> 16, TKi = packed & 0xffff;
    const int T = TJi * M.ntk + TKi;
    const int jo = TJi * TJ + L.jl, ko = TKi * TK + L.kl;
    L.col = (jo < g.ny) && (ko < g.nz);
    const int joc = jo < g.ny ? jo : g.ny - 1, koc = ko < g.nz ? ko : g.nz - 1;
    c0 = ix(g, 0, 2 + joc, 2 + koc);
    int Tb, Tc;
    if (FWD) {
        L.hasb = L.jl >= 1;
        L.hasc = L.kl >= 1;
        L.need_b = L.jl == 0 && TJi >= 1;
        L.need_c = L.kl == 0 && TKi >= 1;
        L.out_b = L.jl == TJ - 1;
        L.out_c = L.kl == TK - 1;
        L.ib = L.hasb ? L.tid - TK : TJ * TK;
        L.ic = L.hasc ? L.tid - 1 : TJ * TK;
        Tb = T - M.ntk;
        Tc = T - 1;
        w.dtb = TJ - 1;
        w.dtc = TK - 1;
    } else {
        // a downstream column outside the slab stands for ghost cells: zero
        L.hasb = L.jl <= TJ - 2 && jo + 1 < g.ny;
        L.hasc = L.kl <= TK - 2 && ko + 1 < g.nz;
        L.need_b = L.jl == TJ - 1 && TJi + 1 < M.ntj;
        L.need_c = L.kl == TK - 1 && TKi + 1 < M.ntk;
        L.out_b = L.jl == 0;
        L.out_c = L.kl == 0;
        L.ib = L.hasb ? L.tid + TK : TJ * TK;
        L.ic = L.hasc ? L.tid + 1 : TJ * TK;
        Tb = T + M.ntk;
        Tc = T + 1;
        w.dtb = -(TJ - 1);
        w.dtc = -(TK - 1);
    }
    // record (tile, step t) starts at (tile * S + t) * XR_STEP; the host
    // keeps every offset below XR_OOB
    const unsigned S = (unsigned)M.S;
    L.ob = (unsigned)T * S * XR_STEP + (unsigned)L.kl * XR_ENTRY;
    L.oc = (unsigned)T * S * XR_STEP + (unsigned)L.jl * XR_ENTRY;
    L.pb = L.need_b ? (unsigned)Tb * S * XR_STEP + (unsigned)L.kl * XR_ENTRY : XR_OOB;
    L.pc = L.need_c ? (unsigned)Tc * S * XR_STEP + (unsigned)L.jl * XR_ENTRY : XR_OOB;
    w.g = g;
    w.S = M.S;
    w.K = K;
    w.iK = 1.0 / K;
    w.e = e;
    w.deltaT = deltaT;
    w.h[0] = hx;
    w.h[1] = hy;
    w.h[2] = hz;
    constexpr int F = 0x00020000;   // raw buffer, 32-bit data format
    const long row0 = march_row(M, T, 0);
    w.grow = __builtin_amdgcn_make_buffer_rsrc((void *)(Gm + row0 * G4P * 64), (short)0, -1, F);
    w.urow = __builtin_amdgcn_make_buffer_rsrc((void *)(Um + row0 * F4P * 64), (short)0, -1, F);
    w.rrow = __builtin_amdgcn_make_buffer_rsrc((void *)(Rm + row0 * F4P * 64), (short)0, -1, F);
    w.xrow0 = reinterpret_cast<char *>(Xm + row0 * F4P * 64);
    w.rb = __builtin_amdgcn_make_buffer_rsrc((void *)rec, (short)0, (int)rec_bytes, F);
    w.rc = __builtin_amdgcn_make_buffer_rsrc((void *)(reinterpret_cast<char *>(rec) + rec_bytes), (short)0,
                                             (int)rec_bytes, F);
    w.max_spin = max_spin;
    return T;
}

// the march start: wait until the upstream tiles have published the records
// of step `slack` (so that the polls issued two steps ahead find theirs
// written), the x product of the ghost cell upstream in i, and the polls of
// steps 0 and 1
template <bool FWD>
__device__ __forceinline__ void fused_start(const M4Args &w, const F4Lane &L, int slack, C5 U_,
                                            const double *__restrict__ J, M9 met, C5 Xg, long c0, double sa[5],
                                            unsigned long long nb0[6], unsigned long long nc0[6],
                                            unsigned long long nb1[6], unsigned long long nc1[6], bool &late)
{
    const Geo &g = w.g;
    const int S = w.S;
    if (slack > 0 && slack < S) {
        const int ts = FWD ? slack : S - 1 - slack;
        const int ils = ts - L.jl - L.kl;
        const bool acts = L.col && ils >= 0 && ils < g.nx;
        unsigned long long sb[6], sc[6];
        double tmp[5] = {0, 0, 0, 0, 0};
        const unsigned vb = rec_poll(L.pb, acts && L.need_b, ts, w.dtb);
        const unsigned vc = rec_poll(L.pc, acts && L.need_c, ts, w.dtc);
        rec_issue(w.rb, vb, sb);
        rec_issue(w.rc, vc, sc);
        rec_settle(w.rb, vb, acts && L.need_b, sb, tmp, w.max_spin, late);
        rec_settle(w.rc, vc, acts && L.need_c, sc, tmp, w.max_spin, late);
    }
#pragma unroll
    for (int v = 0; v < 5; v++) sa[v] = 0.0;
    const long sxs = (long)g.Ym * g.Zm;
    if (L.col) ghost_product4<FWD>(w, U_, J, met, Xg, c0 + (long)(FWD ? 2 : g.gend + 1) * sxs, sa);
#pragma unroll
    for (int s0 = 0; s0 < 2; s0++) {
        const int t0 = FWD ? s0 : S - 1 - s0;
        const int il0 = t0 - L.jl - L.kl;
        const bool act0 = s0 < S && L.col && il0 >= 0 && il0 < g.nx;
        rec_issue(w.rb, rec_poll(L.pb, act0 && L.need_b, t0, w.dtb), s0 ? nb1 : nb0);
        rec_issue(w.rc, rec_poll(L.pc, act0 && L.need_c, t0, w.dtc), s0 ? nc1 : nc0);
    }
}

template <int TJ, int TK, bool FWD, bool GC>
__global__ void __launch_bounds__(TJ *TK) k_lusgs_fused(Geo g, March M, double K, double e, double deltaT, double hx,
                                                      double hy, double hz, const double2 *__restrict__ Gm,
                                                      const double2 *__restrict__ Gc,
                                                      const double2 *__restrict__ Um, const double2 *Rm, double2 *Xm,
                                                      C5 U_, const double *__restrict__ J, M9 met, C5 Xg,
                                                      double2 *rec, unsigned rec_bytes,
                                                      const int *__restrict__ order, int ntiles, int *ticket, int *err,
                                                      int max_spin, int slack)
{
    static_assert(TJ * TK == 64, "one wave per tile");
    __shared__ int s_t;
    // in-tile products by step parity, lane, then the zero entry
    __shared__ double sB[2][5][TJ * TK + 1], sC[2][5][TJ * TK + 1];
    if (threadIdx.x == 0) s_t = atomicAdd(ticket, 1);
    if (threadIdx.x < 10) {
        sB[threadIdx.x / 5][threadIdx.x % 5][TJ * TK] = 0.0;
        sC[threadIdx.x / 5][threadIdx.x % 5][TJ * TK] = 0.0;
    }
    __syncthreads();
    // forward: anti-diagonals ascending; backward: the same list reversed
    const int tk = __builtin_amdgcn_readfirstlane(s_t);
    const int packed = order[FWD ? tk : ntiles - 1 - tk];
    F4Lane L;
    M4Args w;
    long c0;
    const int T = fused_setup<TJ, TK, FWD>(L, w, c0, threadIdx.x, packed, g, M, K, e, deltaT, hx, hy, hz, Gm, Um, Rm,
                                           Xm, rec, rec_bytes, max_spin);
    const int S = M.S;
    bool late = false;
    // exchange polls in flight two steps ahead: one register set per parity
    unsigned long long nb0[6], nc0[6], nb1[6], nc1[6];
    double sa[5];
    fused_start<FWD>(w, L, slack, U_, J, met, Xg, c0, sa, nb0, nc0, nb1, nc1, late);
    if (GC) {
#pragma unroll
        for (int p = 0; p < G4P; p++) w.gcol[p] = Gc[((long)T * G4P + p) * 64 + L.tid];
    }
    Ld4 La, Lb;
    Cf4 C;
    load4<FWD, GC>(La, w, (unsigned)L.tid * 16u, 0);
    load4<FWD, GC>(Lb, w, (unsigned)L.tid * 16u, 1);
    coef4<FWD>(C, La, w);
    // nothing outstanding on entry to the march: the compiler's wait counts
    // inside the loop then follow the loop's own loads and stores (merged
    // with the prologue's, they were full waits at every step)
    __builtin_amdgcn_s_waitcnt(0);
    for (int s = 0; s < S; s += 2) {
        step4<TJ, TK, FWD, GC, 0>(w, L, s, C, Lb, La, sa, sB, sC, nb0, nc0, late);
        if (s + 1 < S) step4<TJ, TK, FWD, GC, 1>(w, L, s + 1, C, La, Lb, sa, sB, sC, nb1, nc1, late);
    }
    if (late) atomicOr(err, 1);
}

// ---------------------------------------------------------------------------
// Schedule 4 with helper waves.  The march step of k_lusgs_fused spends about
// half its time on the coefficients of the next cell, work that does not
// depend on the sweep.  Here a tile's workgroup holds one chain wave and H
// helper waves.  Helper h loads the state of steps h, h + H, h + 2H, ...,
// computes each cell's compact coefficient set (coef_compact) and leaves it,
// with the residual, in one of R LDS slots; the chain wave runs the march as
// k_lusgs_fused does, taking each step's coefficients from its slot and
// forming the metric products itself, so the increments are the same bits.
// Slot protocol (all in LDS, whose operations from one wave complete in
// issue order): a helper writes slot s % R once the chain wave has posted
// `consumed` > s - R, then sets ready[s % R] = s; the chain wave posts
// consumed = s when it starts step s (the slots of earlier steps have been
// read by then) and reads slot s % R after seeing ready[s % R] == s.  Both
// waits are bounded spins that set the error flag on expiry.
#ifndef FH_R
#define FH_R 3                 // LDS coefficient slots per tile
#endif
constexpr int FH_NP_GC = 13;   // slot pairs per lane: 21 coefficients + 5 residuals (+ 1 pad)
constexpr int FH_NP = 18;      // and the nine metrics where they vary along i

__device__ __forceinline__ void lds_fence() { asm volatile("" ::: "memory"); }
// relaxed workgroup-scope accesses to the slot flags: single LDS operations
// with no wait attached (ordering comes from LDS completing a wave's
// operations in issue order, and lds_fence keeps the compiler's order)
__device__ __forceinline__ int flag_get(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void flag_put(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool FWD, bool GC, int R, int H, int NP>
__device__ void fused_helper(const M4Args &w, int h, int tid, double2 (*slot)[NP][64], int *ready, int *consumed,
                             bool &late)
{
    const int S = w.S;
    const unsigned vo = (unsigned)tid * 16u;
    Ld4 A, B;
    load4<FWD, GC>(A, w, vo, h);
    auto one = [&](int s, const Ld4 &L) {
        const double m[9] = {L.g[0].y, L.g[1].x, L.g[1].y, L.g[2].x, L.g[2].y, L.g[3].x, L.g[3].y, L.g[4].x, L.g[4].y};
        const double U[5] = {L.u[0].x, L.u[0].y, L.u[1].x, L.u[1].y, L.u[2].x};
        double c[CF_N];
        coef_compact<FWD>(c, L.g[0].x, m, U, w);
        const int need = s - R + 1;   // the chain wave has started step s - R + 1
        if (need > 0 && __builtin_amdgcn_readfirstlane(flag_get(consumed)) < need) {
            int n = 0;
            while (__builtin_amdgcn_readfirstlane(flag_get(consumed)) < need && n < w.max_spin) {
                __builtin_amdgcn_s_sleep(1);
                n++;
            }
            late |= n >= w.max_spin;
        }
        lds_fence();
        double2 *o = &slot[s % R][0][tid];
#pragma unroll
        for (int p = 0; p < 10; p++) o[p * 64] = make_double2(c[2 * p], c[2 * p + 1]);
        o[10 * 64] = make_double2(c[20], L.r[0].x);
        o[11 * 64] = make_double2(L.r[0].y, L.r[1].x);
        o[12 * 64] = make_double2(L.r[1].y, L.r[2].x);
        if constexpr (!GC) {
#pragma unroll
            for (int p = 0; p < 4; p++) o[(13 + p) * 64] = make_double2(m[2 * p], m[2 * p + 1]);
            o[17 * 64] = make_double2(m[8], 0.0);
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        flag_put(ready + s % R, s);
        lds_fence();
    };
    for (int s = h; s < S; s += 2 * H) {
        load4<FWD, GC>(B, w, vo, s + H);
        one(s, A);
        if (s + H >= S) break;
        load4<FWD, GC>(A, w, vo, s + 2 * H);
        one(s + H, B);
    }
}

// one march step of the chain wave: step4 with the coefficients from the
// step's slot instead of computed, and no workgroup barrier (the chain
// wave's own LDS operations complete in issue order)
template <int TJ, int TK, bool FWD, bool GC, int PAR, int R, int NP>
__device__ __forceinline__ void step_chain(const M4Args &w, const F4Lane &L, int s, const double mc[9],
                                           const double2 (*slot)[NP][64], int *ready, int *consumed,
                                           double sa[5], double (*sB)[5][TJ * TK + 1],
                                           double (*sC)[5][TJ * TK + 1], unsigned long long nb[6],
                                           unsigned long long nc[6], bool &late)
{
    const int t = FWD ? s : w.S - 1 - s;
    const int il = t - L.jl - L.kl;
    const bool act = L.col && il >= 0 && il < w.g.nx;
    const int r = s % R;
    lds_fence();
    flag_put(consumed, s);
    const int rd = flag_get(ready + r);
    lds_fence();
    double2 q[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) q[p] = slot[r][p][L.tid];
    double yb[5], yc[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        yb[v] = sB[PAR ^ 1][v][L.ib];
        yc[v] = sC[PAR ^ 1][v][L.ic];
    }
    if (__builtin_amdgcn_readfirstlane(rd) != s) {
        int n = 0;
        while (__builtin_amdgcn_readfirstlane(flag_get(ready + r)) != s && n < w.max_spin) {
            __builtin_amdgcn_s_sleep(1);
            n++;
        }
        late |= n >= w.max_spin;
        lds_fence();
#pragma unroll
        for (int p = 0; p < NP; p++) q[p] = slot[r][p][L.tid];
    }
    rec_settle2(w, rec_poll(L.pb, act && L.need_b, t, w.dtb), rec_poll(L.pc, act && L.need_c, t, w.dtc), nb, nc,
                yb, yc, late);
    double c[CF_N], rr[5], m[9];
#pragma unroll
    for (int p = 0; p < 10; p++) {
        c[2 * p] = q[p].x;
        c[2 * p + 1] = q[p].y;
    }
    c[20] = q[10].x;
    rr[0] = q[10].y;
    rr[1] = q[11].x;
    rr[2] = q[11].y;
    rr[3] = q[12].x;
    rr[4] = q[12].y;
    if constexpr (GC) {
#pragma unroll
        for (int v = 0; v < 9; v++) m[v] = mc[v];
    } else {
#pragma unroll
        for (int v = 0; v < 9; v++) m[v] = v & 1 ? q[13 + v / 2].y : q[13 + v / 2].x;
    }
    const double d11 = c[18], d22 = c[19], d51 = c[20];
    double x[5];
    if (FWD) {
        const double LL0 = sa[0] + yb[0] + yc[0], LL1 = sa[1] + yb[1] + yc[1], LL2 = sa[2] + yb[2] + yc[2];
        const double LL3 = sa[3] + yb[3] + yc[3], LL4 = sa[4] + yb[4] + yc[4];
        x[0] = d11 * (rr[0] - LL0);
        x[1] = d22 * (rr[1] - LL1);
        x[2] = d22 * (rr[2] - LL2);
        x[3] = d22 * (rr[3] - LL3);
        x[4] = d51 * (rr[0] - LL0) + d22 * (rr[4] - LL4);
    } else {
        const double UU0 = sa[0] + yb[0] + yc[0], UU1 = sa[1] + yb[1] + yc[1], UU2 = sa[2] + yb[2] + yc[2];
        const double UU3 = sa[3] + yb[3] + yc[3], UU4 = sa[4] + yb[4] + yc[4];
        x[0] = rr[0] - d11 * UU0;
        x[1] = rr[1] - d22 * UU1;
        x[2] = rr[2] - d22 * UU2;
        x[3] = rr[3] - d22 * UU3;
        x[4] = rr[4] - (d51 * UU0 + d22 * UU4);
    }
    double pa[5], pb[5], pc[5];
    prod_c<FWD>(c, m, x, pa);
    prod_c<FWD>(c + 6, m + 3, x, pb);
    prod_c<FWD>(c + 12, m + 6, x, pc);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        sB[PAR][v][L.tid] = pb[v];
        sC[PAR][v][L.tid] = pc[v];
        sa[v] = act ? pa[v] : sa[v];
    }
    {   // the exchange polls two steps ahead
        const int tn = FWD ? s + 2 : w.S - 3 - s;
        const int iln = tn - L.jl - L.kl;
        const bool actn = s + 2 < w.S && L.col && iln >= 0 && iln < w.g.nx;
        rec_issue(w.rb, rec_poll(L.pb, actn && L.need_b, tn, w.dtb), nb);
        rec_issue(w.rc, rec_poll(L.pc, actn && L.need_c, tn, w.dtc), nc);
    }
    __builtin_amdgcn_sched_barrier(0);   // loads before stores
    double2 *xo = reinterpret_cast<double2 *>(w.xrow0 + (long)t * F4ROWB) + L.tid;
    xo[0] = make_double2(x[0], x[1]);
    xo[64] = make_double2(x[2], x[3]);
    xo[128] = make_double2(x[4], 0.0);
    rec_store(w.rb, rec_pub(L.ob, act && L.out_b, t), pb);
    rec_store(w.rc, rec_pub(L.oc, act && L.out_c, t), pc);
}

template <int TJ, int TK, bool FWD, bool GC, int R, int H>
__global__ void __launch_bounds__(64 * (1 + H)) k_lusgs_helped(Geo g, March M, double K, double e, double deltaT,
                                                             double hx, double hy, double hz,
                                                             const double2 *__restrict__ Gm,
                                                             const double2 *__restrict__ Gc,
                                                             const double2 *__restrict__ Um, const double2 *Rm,
                                                             double2 *Xm, C5 U_, const double *__restrict__ J, M9 met,
                                                             C5 Xg, double2 *rec, unsigned rec_bytes,
                                                             const int *__restrict__ order, int ntiles, int *ticket,
                                                             int *err, int max_spin, int slack)
{
    static_assert(TJ * TK == 64, "one chain wave per tile");
    constexpr int NP = GC ? FH_NP_GC : FH_NP;
    __shared__ int s_t;
    __shared__ int s_ready[R], s_consumed;
    __shared__ double sB[2][5][TJ * TK + 1], sC[2][5][TJ * TK + 1];
    __shared__ double2 slot[R][NP][64];
    const int wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        s_t = atomicAdd(ticket, 1);
        s_consumed = 0;
    }
    if (threadIdx.x < R) s_ready[threadIdx.x] = -1;
    if (threadIdx.x < 10) {
        sB[threadIdx.x / 5][threadIdx.x % 5][TJ * TK] = 0.0;
        sC[threadIdx.x / 5][threadIdx.x % 5][TJ * TK] = 0.0;
    }
    __syncthreads();
    const int tk = __builtin_amdgcn_readfirstlane(s_t);
    const int packed = order[FWD ? tk : ntiles - 1 - tk];
    F4Lane L;
    M4Args w;
    long c0;
    const int T = fused_setup<TJ, TK, FWD>(L, w, c0, threadIdx.x & 63, packed, g, M, K, e, deltaT, hx, hy, hz, Gm, Um,
                                           Rm, Xm, rec, rec_bytes, max_spin);
    const int S = M.S;
    bool late = false;
    if (GC) {
#pragma unroll
        for (int p = 0; p < G4P; p++) w.gcol[p] = Gc[((long)T * G4P + p) * 64 + L.tid];
    }
    if (wv > 0) {
        if (wv - 1 < S) fused_helper<FWD, GC, R, H, NP>(w, wv - 1, L.tid, slot, s_ready, &s_consumed, late);
        if (late) atomicOr(err, 1);
        return;
    }
    unsigned long long nb0[6], nc0[6], nb1[6], nc1[6];
    double sa[5];
    fused_start<FWD>(w, L, slack, U_, J, met, Xg, c0, sa, nb0, nc0, nb1, nc1, late);
    double mc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr (GC) {
        const double gm[9] = {w.gcol[0].y, w.gcol[1].x, w.gcol[1].y, w.gcol[2].x, w.gcol[2].y,
                              w.gcol[3].x, w.gcol[3].y, w.gcol[4].x, w.gcol[4].y};
#pragma unroll
        for (int v = 0; v < 9; v++) mc[v] = gm[v];
    }
    __builtin_amdgcn_s_waitcnt(0);   // nothing outstanding on entry (as in k_lusgs_fused)
    for (int s = 0; s < S; s += 2) {
        step_chain<TJ, TK, FWD, GC, 0, R, NP>(w, L, s, mc, slot, s_ready, &s_consumed, sa, sB, sC, nb0, nc0, late);
        if (s + 1 < S)
            step_chain<TJ, TK, FWD, GC, 1, R, NP>(w, L, s + 1, mc, slot, s_ready, &s_consumed, sa, sB, sC, nb1, nc1,
                                                  late);
    }
    if (late) atomicOr(err, 1);
}

long march4_rec_bytes(const Geo &g)
{
    const March M = march_of(g);
    return (long)M.ntj * M.ntk * M.S * XR_STEP;
}

__global__ void k_fill_rec4(ulonglong2 *rec, long n, int *ticket)
{
    const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) *ticket = 0;
    for (long i = i0; i < n; i += (long)gridDim.x * blockDim.x) rec[i] = make_ulonglong2(SENTINEL, SENTINEL);
}

void launch_rec4_prep(void *rec, long rec_bytes, int *counters, hipStream_t s)
{
    const long n = 2 * rec_bytes / 16;
    const long blocks = std::min((n + 255) / 256, 4096L);
    hipLaunchKernelGGL(k_fill_rec4, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<ulonglong2 *>(rec), n,
                       counters);
}

void launch_gather4(const Geo &g, C5 U_, C5 R, const double *J, M9 met, void *Um, void *Rm, void *Gm, hipStream_t s)
{
    const March M = march_of(g);
    hipLaunchKernelGGL(k_gather4, mgather_grid(M), dim3(256), 0, s, g, M, U_, R, J, met, static_cast<double2 *>(Um),
                       static_cast<double2 *>(Rm), static_cast<double2 *>(Gm));
}

void launch_scatter4(const Geo &g, const void *Xm, F5 NL, F5 NR, hipStream_t s)
{
    const March M = march_of(g);
    hipLaunchKernelGGL(k_scatter4, mgather_grid(M), dim3(256), 0, s, g, M, static_cast<const double2 *>(Xm), NL, NR);
}

void launch_lusgs_fused(const Geo &g, const Phys &p, double deltaT, double e, int forward, const void *Gm,
                        const void *Gc, const void *Um, const void *Rm, void *Xm, C5 U_, const double *J, M9 met, C5 Xg,
                        void *rec, long rec_bytes, const int *order, int ntiles, int *counters,
                        int max_spin, int slack, int waves, hipStream_t s)
{
    const March M = march_of(g);
    const dim3 gr(ntiles), bl(MARCH_TJ * MARCH_TK);
    const double2 *G = static_cast<const double2 *>(Gm), *Gcol = static_cast<const double2 *>(Gc),
                  *Uu = static_cast<const double2 *>(Um), *R = static_cast<const double2 *>(Rm);
    double2 *X = static_cast<double2 *>(Xm);
    const int sl = slack < 0 ? FUSED_SLACK : slack;
    const double hx = 1.0 / p.dxi, hy = 1.0 / p.deta, hz = 1.0 / p.dzeta;
#define CGLI_FUSED(FW, GCOL)                                                                                         \
    hipLaunchKernelGGL((k_lusgs_fused<MARCH_TJ, MARCH_TK, FW, GCOL>), gr, bl, 0, s, g, M, p.K, e, deltaT, hx, hy, hz, \
                       G, Gcol, Uu, R, X, U_, J, met, Xg, static_cast<double2 *>(rec), (unsigned)rec_bytes,     \
                       order, ntiles, counters, counters + 1,                                                        \
                       max_spin, sl)
#define CGLI_HELPED(FW, GCOL, H)                                                                                     \
    hipLaunchKernelGGL((k_lusgs_helped<MARCH_TJ, MARCH_TK, FW, GCOL, FH_R, H>), gr, dim3(64 * (1 + H)), 0, s, g, M, p.K, \
                       e, deltaT, hx, hy, hz, G, Gcol, Uu, R, X, U_, J, met, Xg, static_cast<double2 *>(rec),        \
                       (unsigned)rec_bytes, order, ntiles,                                                           \
                       counters, counters + 1, max_spin, sl)
#define CGLI_HELPED_H(FW, GCOL)                     \
    switch (waves) {                                \
    case 2: CGLI_HELPED(FW, GCOL, 1); break;        \
    case 3: CGLI_HELPED(FW, GCOL, 2); break;        \
    default: CGLI_HELPED(FW, GCOL, 3); break;       \
    }
    if (waves <= 1) {
        if (forward && Gc) CGLI_FUSED(true, true);
        else if (forward) CGLI_FUSED(true, false);
        else if (Gc) CGLI_FUSED(false, true);
        else CGLI_FUSED(false, false);
    } else {
        if (forward && Gc) CGLI_HELPED_H(true, true)
        else if (forward) CGLI_HELPED_H(true, false)
        else if (Gc) CGLI_HELPED_H(false, true)
        else CGLI_HELPED_H(false, false)
    }
#undef CGLI_FUSED
#undef CGLI_HELPED
#undef CGLI_HELPED_H
}

// 1 into *flag where a cell's J or metrics differ from those of the first
// interior plane of its column (interior cells only: the sweep reads no
// other geometry per step)
__global__ void k_geo_icheck(Geo g, const double *__restrict__ J, M9 met, int *flag)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 4 + blockIdx.z;
    if (k > g.NZ || j > g.NY || i > g.gend) return;
    const long c = ix(g, i, j, k), c3 = ix(g, 3, j, k);
    bool same = J[c] == J[c3];
#pragma unroll
    for (int q = 0; q < 9; q++) same = same && met.m[q][c] == met.m[q][c3];
    if (!same) *flag = 1;
}

// the column geometry of every tile lane (plane i = 3), for the GC sweep
__global__ void k_gather_cols4(Geo g, March M, const double *__restrict__ J, M9 met, double2 *__restrict__ Gc)
{
    const int T = blockIdx.x, l = threadIdx.x;
    const int TJi = T / M.ntk, TKi = T % M.ntk;
    const int jo = TJi * MARCH_TJ + l / MARCH_TK, ko = TKi * MARCH_TK + l % MARCH_TK;
    const int joc = jo < g.ny ? jo : g.ny - 1, koc = ko < g.nz ? ko : g.nz - 1;
    const long c = ix(g, 3, 2 + joc, 2 + koc);
    double2 *o = Gc + (long)T * G4P * 64 + l;
    o[0] = make_double2(J[c], met.m[0][c]);
    o[64] = make_double2(met.m[1][c], met.m[2][c]);
    o[128] = make_double2(met.m[3][c], met.m[4][c]);
    o[192] = make_double2(met.m[5][c], met.m[6][c]);
    o[256] = make_double2(met.m[7][c], met.m[8][c]);
}

void launch_geo_icheck(const Geo &g, const double *J, M9 met, int *flag, hipStream_t s)
{
    if (g.nx < 2) return;
    const dim3 b(64, 4);
    const dim3 gr((g.nz + b.x - 1) / b.x, (g.ny + b.y - 1) / b.y, g.nx - 1);
    hipLaunchKernelGGL(k_geo_icheck, gr, b, 0, s, g, J, met, flag);
}

long march4_col_pairs(const Geo &g)
{
    const March M = march_of(g);
    return (long)M.ntj * M.ntk * G4P * 64;
}

void launch_gather_cols4(const Geo &g, const double *J, M9 met, void *Gc, hipStream_t s)
{
    const March M = march_of(g);
    hipLaunchKernelGGL(k_gather_cols4, dim3(M.ntj * M.ntk), dim3(64), 0, s, g, M, J, met, static_cast<double2 *>(Gc));
}

// the launchers take the allocations (march_*_pairs / march_field_doubles)
void launch_split_march(const Geo &g, const Phys &p, double deltaT, double e, C5 U_, const double *J, M9 met,
                        C5 MR, void *coefM, double *dgnM, double *mrM, int fp32, hipStream_t s)
{
    const March M = march_of(g);
    const dim3 gr = mgather_grid(M), bl(256);
    if (fp32)
        hipLaunchKernelGGL(k_split_march<float>, gr, bl, 0, s, g, p, deltaT, e, U_, J, met, MR, M,
                           static_cast<float *>(coefM), dgnM, mrM);
    else
        hipLaunchKernelGGL(k_split_march<double>, gr, bl, 0, s, g, p, deltaT, e, U_, J, met, MR, M,
                           static_cast<double *>(coefM), dgnM, mrM);
}

void launch_march_scatter(const Geo &g, const double *nrM, F5 NL, F5 NR, hipStream_t s)
{
    const March M = march_of(g);
    hipLaunchKernelGGL(k_march_scatter, mgather_grid(M), dim3(256), 0, s, g, M, nrM, NL, NR);
}

void launch_lusgs_march(const Geo &g, int forward, const void *coefM, const double *dgnM, int fp32,
                        const double *Rm, double *Xm, C5 Xg, double *xbuf, double *trash, const int *order,
                        int ntiles, int *counters, int max_spin, int slack, hipStream_t s)
{
    const March M = march_of(g);
    const dim3 gr(ntiles), bl(MARCH_TJ * MARCH_TK);
    const float *cf = static_cast<const float *>(coefM);
    const double *cd = static_cast<const double *>(coefM);
    if (forward && fp32)
        hipLaunchKernelGGL((k_lusgs_march<MARCH_TJ, MARCH_TK, true, float>), gr, bl, 0, s, g, M, cf, dgnM, Rm, Xm, Xg,
                           xbuf, trash, order, ntiles, counters, counters + 1, max_spin, slack);
    else if (forward)
        hipLaunchKernelGGL((k_lusgs_march<MARCH_TJ, MARCH_TK, true, double>), gr, bl, 0, s, g, M, cd, dgnM, Rm, Xm, Xg,
                           xbuf, trash, order, ntiles, counters, counters + 1, max_spin, slack);
    else if (fp32)
        hipLaunchKernelGGL((k_lusgs_march<MARCH_TJ, MARCH_TK, false, float>), gr, bl, 0, s, g, M, cf, dgnM, Rm, Xm, Xg,
                           xbuf, trash, order, ntiles, counters, counters + 1, max_spin, slack);
    else
        hipLaunchKernelGGL((k_lusgs_march<MARCH_TJ, MARCH_TK, false, double>), gr, bl, 0, s, g, M, cd, dgnM, Rm, Xm, Xg,
                           xbuf, trash, order, ntiles, counters, counters + 1, max_spin, slack);
}

// tile order: anti-diagonals J+K ascending, J ascending inside
static std::vector<int> tile_order(int ntj, int ntk)
{
    std::vector<int> o;
    for (int s = 0; s <= ntj + ntk - 2; s++)
        for (int J = 0; J < ntj; J++) {
            const int K = s - J;
            if (K >= 0 && K < ntk) o.push_back((J << 16) | K);
        }
    return o;
}

// split blocks sized for FP64 elements (an FP32 sweep uses half), and the
// diagonal blocks
long coef_doubles(const Geo &g) { return (long)(g.nx + 6) * g.Ym * coef_nkb(g) * CBLK; }
long diag_doubles(const Geo &g) { return (long)(g.nx + 6) * g.Ym * coef_nkb(g) * DBLK; }

int wave_tiles(const Geo &g, int tj, int tk, std::vector<int> &order)
{
    const int ntj = (g.ny + tj - 1) / tj, ntk = (g.nz + tk - 1) / tk;
    order = tile_order(ntj, ntk);
    return ntj * ntk;
}

void launch_split_coef(const Geo &g, const Phys &p, double deltaT, double e, C5 U_, const double *J, M9 met,
                       void *coef, double *dgn, int fp32, hipStream_t s)
{
    dim3 b(64, 4);
    dim3 gr((g.NZZ + b.x - 1) / b.x, (g.NYY + b.y - 1) / b.y, g.gend);   // i=2..gend+1
    if (fp32)
        hipLaunchKernelGGL(k_split_coef<float>, gr, b, 0, s, g, p, deltaT, e, U_, J, met,
                           static_cast<float *>(coef), dgn);
    else
        hipLaunchKernelGGL(k_split_coef<double>, gr, b, 0, s, g, p, deltaT, e, U_, J, met,
                           static_cast<double *>(coef), dgn);
}

void launch_march_prep(const Geo &g, int forward, double *xbuf, int *counters, hipStream_t s, double *xcol)
{
    const int rj = forward ? MARCH_TJ - 1 : 0, rk = forward ? MARCH_TK - 1 : 0;
    const long nrow = rj < g.ny ? (g.ny - 1 - rj) / MARCH_TJ + 1 : 0;
    const long ncol = rk < g.nz ? (g.nz - 1 - rk) / MARCH_TK + 1 : 0;
    const long n = std::max(std::max(5L * g.nz * nrow, 5L * g.ny * ncol), 1L);
    const dim3 gr((unsigned)((n + 255) / 256), 2, g.nx);
    hipLaunchKernelGGL(k_fill_sentinel_edges, gr, dim3(256), 0, s, g, xbuf, xcol ? xcol : xbuf, MARCH_TJ, MARCH_TK,
                       rj, rk, counters);
}

void launch_lusgs_wave_prep(const Geo &g, double *xbuf, int *counters, hipStream_t s)
{
    // sentinels into the exchange buffer; counters[0] (ticket) back to 0
    dim3 bf(64, 4);
    dim3 gf((g.NZ - 1 + bf.x - 1) / bf.x, (g.NY - 1 + bf.y - 1) / bf.y, g.nx);
    hipLaunchKernelGGL(k_fill_sentinel, gf, bf, 0, s, g, xbuf);
    hipLaunchKernelGGL(k_reset_ints, dim3(1), dim3(64), 0, s, counters, 1);
}

void launch_lusgs_wave(const Geo &g, int forward, const void *coef, const double *dgn, int fp32, C5 R, F5 X,
                       double *xbuf, double *trash, const int *order, int ntiles, int *counters, int max_spin,
                       hipStream_t s)
{
    // counters[0]: ticket, counters[1]: error flag (cleared by the caller)
    const dim3 gr(ntiles), bl(WAVE_TJ * WAVE_TK);
    const float *cf = static_cast<const float *>(coef);
    const double *cd = static_cast<const double *>(coef);
    if (forward && fp32)
        hipLaunchKernelGGL((k_lusgs_wave<WAVE_TJ, WAVE_TK, true, float>), gr, bl, 0, s, g, cf, dgn, R, X, xbuf, trash,
                           order, counters, counters + 1, max_spin);
    else if (forward)
        hipLaunchKernelGGL((k_lusgs_wave<WAVE_TJ, WAVE_TK, true, double>), gr, bl, 0, s, g, cd, dgn, R, X, xbuf,
                           trash, order, counters, counters + 1, max_spin);
    else if (fp32)
        hipLaunchKernelGGL((k_lusgs_wave<WAVE_TJ, WAVE_TK, false, float>), gr, bl, 0, s, g, cf, dgn, R, X, xbuf,
                           trash, order, counters, counters + 1, max_spin);
    else
        hipLaunchKernelGGL((k_lusgs_wave<WAVE_TJ, WAVE_TK, false, double>), gr, bl, 0, s, g, cd, dgn, R, X, xbuf,
                           trash, order, counters, counters + 1, max_spin);
}

void launch_reset_ints(int *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_ints, dim3(1), dim3(64), 0, s, p, n);
}

}  // namespace cgli
