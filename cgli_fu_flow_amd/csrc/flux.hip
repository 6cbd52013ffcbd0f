// flux.hip -- inviscid face fluxes on gfx950 (FP64).
//
// One thread per face: the thread rebuilds the J-weighted MUSCL-5 left and
// right states of its face from the six cells along the sweep direction
// (Flux_X.cpp:106-216, Flux_Y.cpp:106-249, Flux_Z.cpp:105-216) and applies
// the low-Mach preconditioned Roe flux (Flux_X.cpp:229-350).  The
// reference's ML/MR scratch never goes to HBM.  Operation order follows the
// reference expression by expression and the library is compiled without
// FMA contraction, so results match the CPU build to the last bit except
// where libm and the device math library round sqrt/pow differently.
#include "cgli_internal.hpp"

namespace cgli {

// J-weighted reconstruction weights.  `temp` is evaluated once per stencil
// and shared by the five variables, as in the reference loop body.
struct Rec {
    int kind;     // 5: 5-point left, -5: 5-point right, 3: 3rd left, -3: 3rd right, 2: average
    long a, b, c, d, f;
    double w;
};

__device__ __forceinline__ Rec rec5L(const double *J, long a, long b, long c, long d, long f)
{
    Rec r{5, a, b, c, d, f, 0.0};
    r.w = 1. / (2 / J[a] - 13 / J[b] + 47 / J[c] + 27 / J[d] - 3 / J[f]);
    return r;
}
__device__ __forceinline__ Rec rec5R(const double *J, long a, long b, long c, long d, long f)
{
    Rec r{-5, a, b, c, d, f, 0.0};
    r.w = 1. / (-3 / J[a] + 27 / J[b] + 47 / J[c] - 13 / J[d] + 2 / J[f]);
    return r;
}
__device__ __forceinline__ Rec rec3L(const double *J, long m, long c, long p)
{
    Rec r{3, m, c, p, 0, 0, 0.0};
    r.w = 1. / (-1 / J[m] + 5 / J[c] + 2 / J[p]);
    return r;
}
__device__ __forceinline__ Rec rec3R(const double *J, long m, long c, long p)
{
    Rec r{-3, m, c, p, 0, 0, 0.0};
    r.w = 1. / (2 / J[m] + 5 / J[c] - 1 / J[p]);
    return r;
}
__device__ __forceinline__ Rec rec2(const double *J, long a, long b)
{
    Rec r{2, a, b, 0, 0, 0, 0.0};
    r.w = 1. / (1 / J[a] + 1 / J[b]);
    return r;
}

__device__ __forceinline__ double apply(const Rec &r, const double *__restrict__ U)
{
    switch (r.kind) {
    case 5: return r.w * (2 * U[r.a] - 13 * U[r.b] + 47 * U[r.c] + 27 * U[r.d] - 3 * U[r.f]);
    case -5: return r.w * (-3 * U[r.a] + 27 * U[r.b] + 47 * U[r.c] - 13 * U[r.d] + 2 * U[r.f]);
    case 3: return r.w * (-U[r.a] + 5 * U[r.b] + 2 * U[r.c]);
    case -3: return r.w * (2 * U[r.a] + 5 * U[r.b] - U[r.c]);
    default: return r.w * (U[r.a] + U[r.b]);
    }
}

// Preconditioned Roe flux through one face (Flux_X.cpp:251-350).  (mx,my,mz)
// are the face metrics, (nx_,ny_,nz_) the vector multiplying the pressure
// jump in the momentum dissipation.
__device__ __forceinline__ void roe_face(const double L[5], const double Rq[5], double mx, double my, double mz,
                                         double Jf, double e, double K, double nx_, double ny_, double nz_,
                                         double F[5], double &pL_out, double &pR_out, double &Fav3_out)
{
    const double insqr = 1.0 / sqrt(mx * mx + my * my + mz * mz);

    const double rL = L[0];
    const double uL = L[1] / rL, vL = L[2] / rL, wL = L[3] / rL;
    const double UnL = mx * uL + my * vL + mz * wL;
    const double qL = uL * uL + vL * vL + wL * wL;
    const double pL = (L[4] - 0.5 * rL * qL) * (K - 1);
    const double cL = K * pL / rL;
    const double hL = 0.5 * qL + cL / (K - 1);

    const double rR = Rq[0];
    const double uR = Rq[1] / rR, vR = Rq[2] / rR, wR = Rq[3] / rR;
    const double UnR = mx * uR + my * vR + mz * wR;
    const double qR = uR * uR + vR * vR + wR * wR;
    const double pR = (Rq[4] - 0.5 * rR * qR) * (K - 1);
    const double cR = K * pR / rR;
    const double hR = 0.5 * qR + cR / (K - 1);

    const double sL = sqrt(rL), sR = sqrt(rR);
    const double ss = sL + sR;
    const double rho = sqrt(rL * rR);
    const double u = (sL * uL + sR * uR) / ss;
    const double v = (sL * vL + sR * vR) / ss;
    const double w = (sL * wL + sR * wR) / ss;
    const double Un = (sL * UnL + sR * UnR) / ss;
    const double q = u * u + v * v + w * w;
    const double H = (sL * hL + sR * hR) / ss;
    const double C = (H - 0.5 * q) * (K - 1);

    const double dU1 = rR - rL;
    const double dU2 = rR * uR - rL * uL;
    const double dU3 = rR * vR - rL * vL;
    const double dU4 = rR * wR - rL * wL;
    const double dU5 = (pR / (K - 1) + 0.5 * rR * qR) - (pL / (K - 1) + 0.5 * rL * qL);

    const double qc = q / C;
    const double beta = (qc > e) ? qc : e;
    const double ubar = 0.5 * (1 + beta) * Un * insqr;
    const double S = 0.5 * sqrt(4 * beta * C + Un * insqr * Un * insqr * (1 - beta) * (1 - beta));
    const double th = q / C;
    const double U_p = 0.5 * (1 + th) * Un * insqr;
    const double C_p = 0.5 * sqrt(4 * C * th + Un * insqr * Un * insqr * (1 - th) * (1 - th));

    const double t1 = (pR - pL) / rho / beta / C;
    const double t2 = U_p / S * (UnR - UnL) * insqr;
    const double t3 = 0.5 * (1 - beta) * Un * insqr * ubar / S;
    const double t4 = 0.5 * (1 - th) * Un * insqr * U_p / S;
    const double dUc = (S - t3 - beta * fabs(Un) * insqr) * t1 + t2;
    const double dPc = U_p / S * (pR - pL) + (C_p - fabs(Un) * insqr + t4) * rho * (UnR - UnL) * insqr;

    const double aU = fabs(Un * insqr);
    const double Fav1 = aU * dU1 + dUc * rho;
    const double Fav2 = aU * dU2 + dUc * rho * u + dPc * nx_;
    const double Fav3 = aU * dU3 + dUc * rho * v + dPc * ny_;
    const double Fav4 = aU * dU4 + dUc * rho * w + dPc * nz_;
    const double Fav5 = aU * dU5 + dUc * rho * H + dPc * Un * insqr;

    F[0] = 0.5 * ((rL * UnL + rR * UnR) - Fav1 / insqr) / Jf;
    F[1] = 0.5 * ((rL * uL * UnL + rR * uR * UnR + mx * (pL + pR)) - Fav2 / insqr) / Jf;
    F[2] = 0.5 * ((rL * vL * UnL + rR * vR * UnR + my * (pL + pR)) - Fav3 / insqr) / Jf;
    F[3] = 0.5 * ((rL * wL * UnL + rR * wR * UnR + mz * (pL + pR)) - Fav4 / insqr) / Jf;
    F[4] = 0.5 * ((UnL * (3.5 * pL + 0.5 * rL * qL) + UnR * (3.5 * pR + 0.5 * rR * qR)) - Fav5 / insqr) / Jf;
    pL_out = pL;
    pR_out = pR;
    Fav3_out = Fav3;
}

// ---------------------------------------------------------------- x faces
// Face i+1/2 between cells (i, j+1, k+1) and (i+1, j+1, k+1), stored at
// [i][j][k] for i=2..gend, j=1..NY-1, k=1..NZ-1.
__global__ void __launch_bounds__(256) k_flux_x(Geo g, double K, double e, C5 U, const double *__restrict__ J,
                                                const double *__restrict__ Jf, M9 mf, F5 F)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 2 + blockIdx.z;
    if (k > g.NZ - 1 || j > g.NY - 1) return;
    const int cj = j + 1, ck = k + 1;
    const long sx = g.Ym * g.Zm;
    const long c0 = ix(g, i, cj, ck);
    // left state of face i+1/2; later reference writes win (last-rank patch
    // after first-rank patch after the 5-point loop)
    Rec L;
    if (g.last && i == g.gend) L = rec3L(J, c0 - sx, c0, c0 + sx);
    else if (g.first && i == 3) L = rec3L(J, c0 - sx, c0, c0 + sx);
    else if (g.first && i == 2) L = rec2(J, c0, c0 + sx);
    else L = rec5L(J, c0 - 2 * sx, c0 - sx, c0, c0 + sx, c0 + 2 * sx);
    // right state of face i+1/2 (from cell i+1)
    const long c1 = c0 + sx;
    Rec R;
    if (g.last && i == g.gend) R = rec2(J, c0, c1);
    else if (g.last && i == g.gend - 1) R = rec3R(J, c0, c1, c1 + sx);
    else if (g.first && i == 2) R = rec3R(J, c0, c1, c1 + sx);
    else R = rec5R(J, c1 - 2 * sx, c1 - sx, c1, c1 + sx, c1 + 2 * sx);
    double Lq[5], Rq[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        Lq[v] = apply(L, U.v[v]);
        Rq[v] = apply(R, U.v[v]);
    }
    const long f = ix(g, i, j, k);
    const double xix = mf.m[0][f], xiy = mf.m[1][f], xiz = mf.m[2][f];
    const double insqr = 1.0 / sqrt(xix * xix + xiy * xiy + xiz * xiz);
    double out[5], pl, pr, fav3;
    roe_face(Lq, Rq, xix, xiy, xiz, Jf[f], e, K, xix * insqr, xiy * insqr, xiz * insqr, out, pl, pr, fav3);
#pragma unroll
    for (int v = 0; v < 5; v++) F.v[v][f] = out[v];
}

// ---------------------------------------------------------------- y faces
// Face j+1/2 between cells (i+1, j, k+1) and (i+1, j+1, k+1), stored at
// [i][j][k] for i=2..gend-1, j=1..NY, k=1..NZ-1.  Near-wall order drop of
// Flux_Y.cpp:135-249 and the wall-face override of Flux_Y.cpp:400-408.
__global__ void __launch_bounds__(256) k_flux_y(Geo g, double K, double Ep, double e, int xi_mode, C5 U,
                                                const double *__restrict__ J, const double *__restrict__ Jf,
                                                M9 mf, F5 F)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 2 + blockIdx.z;
    if (k > g.NZ - 1 || j > g.NY) return;
    const long sy = g.Zm;
    const long c0 = ix(g, i + 1, j, k + 1);   // cell j
    const long c1 = c0 + sy;                  // cell j+1
    const int NY = g.NY;
    Rec L;
    if (j == NY) L = rec2(J, c0, c0 - sy);
    else if (j == NY - 1) L = rec3L(J, c0 - sy, c0, c0 + sy);
    else if (j == 3) L = rec3L(J, c0 - sy, c0, c0 + sy);
    else if (j == 2) L = rec2(J, c0, c0 + sy);
    else if (j == 1) L = rec2(J, c0, c0 + sy);
    else L = rec5L(J, c0 - 2 * sy, c0 - sy, c0, c0 + sy, c0 + 2 * sy);
    Rec R;
    if (j == NY) R = rec2(J, c0, c1);
    else if (j == NY - 1) R = rec2(J, c1, c0);
    else if (j == NY - 2) R = rec3R(J, c0, c1, c1 + sy);
    else if (j == 2) R = rec3R(J, c0, c1, c1 + sy);
    else if (j == 1) R = rec2(J, c1, c1 + sy);
    else R = rec5R(J, c1 - 2 * sy, c1 - sy, c1, c1 + sy, c1 + 2 * sy);
    double Lq[5], Rq[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        Lq[v] = apply(L, U.v[v]);
        Rq[v] = apply(R, U.v[v]);
    }
    const long f = ix(g, i, j, k);
    const double etx = mf.m[3][f], ety = mf.m[4][f], etz = mf.m[5][f];
    const double insqr = 1.0 / sqrt(etx * etx + ety * ety + etz * etz);
    double nx_ = 0, ny_ = 0, nz_ = 0;
    if (xi_mode) { nx_ = etx * insqr; ny_ = ety * insqr; nz_ = etz * insqr; }
    double out[5], pl, pr, fav3;
    const double jf = Jf[f];
    roe_face(Lq, Rq, etx, ety, etz, jf, e, K, nx_, ny_, nz_, out, pl, pr, fav3);
    if (j == 1 || j == NY) {
        out[0] = 0;
        out[1] = 0;
        out[2] = 0.5 * (ety * (pl + pr) - Ep * fav3) / jf;
        out[3] = 0;
        out[4] = 0;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) F.v[v][f] = out[v];
}

// ---------------------------------------------------------------- z faces
// Face k+1/2 between cells (i+1, j+1, k) and (i+1, j+1, k+1), stored at
// [i][j][k] for i=2..gend-1, j=1..NY-1, k=1..NZ (Flux_Z.cpp:105-216).
__global__ void __launch_bounds__(256) k_flux_z(Geo g, double K, double e, int xi_mode, C5 U,
                                                const double *__restrict__ J, const double *__restrict__ Jf,
                                                M9 mf, F5 F)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 2 + blockIdx.z;
    if (k > g.NZ || j > g.NY - 1) return;
    const long c0 = ix(g, i + 1, j + 1, k);   // cell k
    const long c1 = c0 + 1;                   // cell k+1
    const int NZ = g.NZ;
    Rec L;
    if (k == NZ) L = rec3L(J, c0 - 1, c0, c0 + 1);
    else if (k == 2) L = rec3L(J, c0 - 1, c0, c0 + 1);
    else if (k == 1) L = rec2(J, c0, c0 + 1);
    else L = rec5L(J, c0 - 2, c0 - 1, c0, c0 + 1, c0 + 2);
    Rec R;
    if (k == NZ) R = rec2(J, c0, c1);
    else if (k == NZ - 1) R = rec3R(J, c0, c1, c1 + 1);
    else if (k == 1) R = rec3R(J, c0, c1, c1 + 1);
    else R = rec5R(J, c1 - 2, c1 - 1, c1, c1 + 1, c1 + 2);
    double Lq[5], Rq[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        Lq[v] = apply(L, U.v[v]);
        Rq[v] = apply(R, U.v[v]);
    }
    const long f = ix(g, i, j, k);
    const double ztx = mf.m[6][f], zty = mf.m[7][f], ztz = mf.m[8][f];
    const double insqr = 1.0 / (sqrt(ztx * ztx + zty * zty + ztz * ztz));
    double nx_ = 0, ny_ = 0, nz_ = 0;
    if (xi_mode) { nx_ = ztx * insqr; ny_ = zty * insqr; nz_ = ztz * insqr; }
    double out[5], pl, pr, fav3;
    roe_face(Lq, Rq, ztx, zty, ztz, Jf[f], e, K, nx_, ny_, nz_, out, pl, pr, fav3);
#pragma unroll
    for (int v = 0; v < 5; v++) F.v[v][f] = out[v];
}

static dim3 grid_for(int nk, int nj, int ni, dim3 b)
{
    return dim3((nk + b.x - 1) / b.x, (nj + b.y - 1) / b.y, ni);
}

void launch_flux_x(const Geo &g, const Phys &p, double e, C5 U, const double *J, const double *Jf, M9 mf,
                   F5 F, hipStream_t s)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_flux_x, grid_for(g.NZ - 1, g.NY - 1, g.gend - 1, b), b, 0, s, g, p.K, e, U, J, Jf,
                       mf, F);
}

void launch_flux_y(const Geo &g, const Phys &p, double Ep, double e, int xi_mode, C5 U, const double *J,
                   const double *Jf, M9 mf, F5 F, hipStream_t s)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_flux_y, grid_for(g.NZ - 1, g.NY, g.gend - 2, b), b, 0, s, g, p.K, Ep, e, xi_mode, U,
                       J, Jf, mf, F);
}

void launch_flux_z(const Geo &g, const Phys &p, double e, int xi_mode, C5 U, const double *J, const double *Jf,
                   M9 mf, F5 F, hipStream_t s)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_flux_z, grid_for(g.NZ, g.NY - 1, g.gend - 2, b), b, 0, s, g, p.K, e, xi_mode, U, J,
                       Jf, mf, F);
}

}  // namespace cgli
