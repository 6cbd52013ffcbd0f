// viscous.hip -- viscous and heat-flux terms on gfx950.  The face pass is
// templated on its arithmetic type (double: the reference path; float: the
// FP32 residual path, loading the FP64 primitives and metrics and storing
// FP64 face fluxes).
//
// Three passes (Viscous_terms.cpp:183-905):
//   1. primitives (u, v, w, T) of every cell the stencil touches, once;
//   2. per cell, the fluxes through its low x, y and z faces (each face only
//      where the divergence uses it), from second-order face gradients with
//      four-point cross averages and Sutherland viscosity;
//   3. vF = dF/dxi + dG/deta + dH/dzeta on interior cells.
// Pass 1 replaces the reference's 16 primitive reconstructions per cell
// with one per cell; the values are the same doubles.
#include "cgli_internal.hpp"

namespace cgli {

__global__ void __launch_bounds__(256) k_prim4(Geo g, double K, double R, C5 U, const double *__restrict__ J,
                                               double *__restrict__ P4)
{
    // box i=2..gend+1, j=1..NYY, k=1..NZZ; (j, k) of a plane flattened, k
    // fastest, so that rows of NZZ cells fill whole waves
    const int nk = g.NZZ;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * g.NYY) return;
    const int k = 1 + t % nk;
    const int j = 1 + t / nk;
    const int i = 2 + blockIdx.z;
    const long c = ix(g, i, j, k);
    const double r = U.v[0][c];
    const double u = U.v[1][c] / r, v = U.v[2][c] / r, w = U.v[3][c] / r;
    const double P = (U.v[4][c] - 0.5 * r * (u * u + v * v + w * w)) * (K - 1) * J[c];
    const double T = P / (r * R) / J[c];
    P4[c] = u;
    P4[c + g.N] = v;
    P4[c + 2 * g.N] = w;
    P4[c + 3 * g.N] = T;
}

template <typename Real>
struct Pt { Real u, v, w, T; };

template <typename Real>
__device__ __forceinline__ Pt<Real> ld(const double *__restrict__ P4, long N, long c)
{
    return Pt<Real>{(Real)P4[c], (Real)P4[c + N], (Real)P4[c + 2 * N], (Real)P4[c + 3 * N]};
}

template <typename Real>
struct D10 { Real u, v, w, u2, v2, w2, uv, vw, uw, T; };

template <typename Real>
__device__ __forceinline__ D10<Real> d2(const Pt<Real> &P, const Pt<Real> &M, Real inv)
{
    D10<Real> d;
    d.u = (P.u - M.u) * inv;
    d.v = (P.v - M.v) * inv;
    d.w = (P.w - M.w) * inv;
    d.u2 = (P.u * P.u - M.u * M.u) * inv;
    d.v2 = (P.v * P.v - M.v * M.v) * inv;
    d.w2 = (P.w * P.w - M.w * M.w) * inv;
    d.uv = (P.u * P.v - M.u * M.v) * inv;
    d.vw = (P.v * P.w - M.v * M.w) * inv;
    d.uw = (P.u * P.w - M.u * M.w) * inv;
    d.T = (P.T - M.T) * inv;
    return d;
}

template <typename Real>
__device__ __forceinline__ D10<Real> d4(const Pt<Real> &A, const Pt<Real> &B, const Pt<Real> &Cc, const Pt<Real> &D,
                                        Real inv)
{
    D10<Real> d;
    d.u = (A.u + B.u - Cc.u - D.u) * inv;
    d.v = (A.v + B.v - Cc.v - D.v) * inv;
    d.w = (A.w + B.w - Cc.w - D.w) * inv;
    d.u2 = (A.u * A.u + B.u * B.u - Cc.u * Cc.u - D.u * D.u) * inv;
    d.v2 = (A.v * A.v + B.v * B.v - Cc.v * Cc.v - D.v * D.v) * inv;
    d.w2 = (A.w * A.w + B.w * B.w - Cc.w * Cc.w - D.w * D.w) * inv;
    d.uv = (A.u * A.v + B.u * B.v - Cc.u * Cc.v - D.u * D.v) * inv;
    d.vw = (A.v * A.w + B.v * B.w - Cc.v * Cc.w - D.v * D.w) * inv;
    d.uw = (A.u * A.w + B.u * B.w - Cc.u * Cc.w - D.u * D.w) * inv;
    d.T = (A.T + B.T - Cc.T - D.T) * inv;
    return d;
}

// metric groups of Viscous_terms.cpp:464-519 (index 1-based like the source)
template <typename Real>
struct Grp { Real a[8], b[8], c[8], d[11], e[11], f[11]; };

template <typename Real>
__device__ __forceinline__ void groups(const Real m[9], Grp<Real> &G)
{
    const Real xix = m[0], xiy = m[1], xiz = m[2], etx = m[3], ety = m[4], etz = m[5], ztx = m[6],
                 zty = m[7], ztz = m[8];
    G.a[1] = Real(4. / 3) * xix * xix + xiy * xiy + xiz * xiz;
    G.a[2] = xix * xix + Real(4. / 3) * xiy * xiy + xiz * xiz;
    G.a[3] = xix * xix + xiy * xiy + Real(4. / 3) * xiz * xiz;
    G.a[4] = xix * xix + xiy * xiy + xiz * xiz;
    G.a[5] = Real(1. / 3) * xix * xiy;
    G.a[6] = Real(1. / 3) * xiy * xiz;
    G.a[7] = Real(1. / 3) * xix * xiz;
    G.b[1] = Real(4. / 3) * etx * etx + ety * ety + etz * etz;
    G.b[2] = etx * etx + Real(4. / 3) * ety * ety + etz * etz;
    G.b[3] = etx * etx + ety * ety + Real(4. / 3) * etz * etz;
    G.b[4] = etx * etx + ety * ety + etz * etz;
    G.b[5] = Real(1. / 3) * etx * ety;
    G.b[6] = Real(1. / 3) * ety * etz;
    G.b[7] = Real(1. / 3) * etx * etz;
    G.c[1] = Real(4. / 3) * ztx * ztx + zty * zty + ztz * ztz;
    G.c[2] = ztx * ztx + Real(4. / 3) * zty * zty + ztz * ztz;
    G.c[3] = ztx * ztx + zty * zty + Real(4. / 3) * ztz * ztz;
    G.c[4] = ztx * ztx + zty * zty + ztz * ztz;
    G.c[5] = Real(1. / 3) * ztx * zty;
    G.c[6] = Real(1. / 3) * zty * ztz;
    G.c[7] = Real(1. / 3) * ztx * ztz;
    G.d[1] = Real(4. / 3) * xix * etx + xiy * ety + xiz * etz;
    G.d[2] = xix * etx + Real(4. / 3) * xiy * ety + xiz * etz;
    G.d[3] = xix * etx + xiy * ety + Real(4. / 3) * xiz * etz;
    G.d[4] = xix * etx + xiy * ety + xiz * etz;
    G.d[5] = xix * ety - Real(2. / 3) * xiy * etx;
    G.d[6] = xix * etz - Real(2. / 3) * xiz * etx;
    G.d[7] = xiy * etx - Real(2. / 3) * xix * ety;
    G.d[8] = xiy * etz - Real(2. / 3) * xiz * ety;
    G.d[9] = xiz * etx - Real(2. / 3) * xix * etz;
    G.d[10] = xiz * ety - Real(2. / 3) * xiy * etz;
    G.e[1] = Real(4. / 3) * xix * ztx + xiy * zty + xiz * ztz;
    G.e[2] = xix * ztx + Real(4. / 3) * xiy * zty + xiz * ztz;
    G.e[3] = xix * ztx + xiy * zty + Real(4. / 3) * xiz * ztz;
    G.e[4] = xix * ztx + xiy * zty + xiz * ztz;
    G.e[5] = xix * zty - Real(2. / 3) * xiy * ztx;
    G.e[6] = xix * ztz - Real(2. / 3) * xiz * ztx;
    G.e[7] = xiy * ztx - Real(2. / 3) * xix * zty;
    G.e[8] = xiy * ztz - Real(2. / 3) * xiz * zty;
    G.e[9] = xiz * ztx - Real(2. / 3) * xix * ztz;
    G.e[10] = xiz * zty - Real(2. / 3) * xiy * ztz;
    G.f[1] = Real(4. / 3) * etx * ztx + ety * zty + etz * ztz;
    G.f[2] = etx * ztx + Real(4. / 3) * ety * zty + etz * ztz;
    G.f[3] = etx * ztx + ety * zty + Real(4. / 3) * etz * ztz;
    G.f[4] = etx * ztx + ety * zty + etz * ztz;
    G.f[5] = etx * zty - Real(2. / 3) * ety * ztx;
    G.f[6] = etx * ztz - Real(2. / 3) * etz * ztx;
    G.f[7] = ety * ztx - Real(2. / 3) * etx * zty;
    G.f[8] = ety * ztz - Real(2. / 3) * etz * zty;
    G.f[9] = etz * ztx - Real(2. / 3) * etx * ztz;
    G.f[10] = etz * zty - Real(2. / 3) * ety * ztz;
}

template <typename Real>
__device__ __forceinline__ Real suth(const Phys &p, Real T)
{
    const Real Ts = (Real)p.suth_T, S = (Real)p.suth_S;
    return (Real)p.mu_L * pow((T / Ts), Real(1.5)) * (Ts + S) / (T + S);
}

// Face fluxes of cell (i,j,k): x face i-1/2 for i<=gend+1 (j<=NY, k<=NZ),
// y face j-1/2 for j<=NYY (i<=gend, k<=NZ), z face k-1/2 for k<=NZZ
// (i<=gend, j<=NY).  Faces go to FF[var*3+dir] (vars 2..5 -> 0..3).
// DIR is the face direction (one launch each): a thread computes one face,
// so it holds one direction's stencil and metric groups (the three
// together needed about 250 VGPRs, two waves per SIMD).
template <typename Real, int DIR>
__global__ void __launch_bounds__(256) k_visc_faces(Geo g, Phys p, const double *__restrict__ P4,
                                                    const double *__restrict__ Ju, M9 mu,
                                                    const double *__restrict__ Jv, M9 mv,
                                                    const double *__restrict__ Jw, M9 mw,
                                                    double *__restrict__ FF)
{
    // (j, k) = (2..NYY, 2..NZZ) of a plane flattened, k fastest
    const int nk = g.NZZ - 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * (g.NYY - 1)) return;
    const int k = 2 + t % nk;
    const int j = 2 + t / nk;
    const int i = 3 + blockIdx.z;
    constexpr int dir = DIR;
    const bool xf = dir == 0 && (j <= g.NY && k <= g.NZ);
    const bool yf = dir == 1 && (i <= g.gend && k <= g.NZ);
    const bool zf = dir == 2 && (i <= g.gend && j <= g.NY);
    if (!(xf || yf || zf)) return;
    const long N = g.N, sx = g.Ym * g.Zm, sy = g.Zm;
    const long c = ix(g, i, j, k);
    const long fid = ix(g, i - 1, j - 1, k - 1);
    const Real invXI = (Real)(1. / (p.dxi)), invET = (Real)(1. / (p.deta)), invZT = (Real)(1. / (p.dzeta));
    const Real inv4XI = (Real)(1. / (4 * p.dxi)), inv4ET = (Real)(1. / (4 * p.deta)),
               inv4ZT = (Real)(1. / (4 * p.dzeta));
    const Real CvK = (Real)(p.Cv * p.K), Pr = (Real)p.Pr_L;
    const Pt<Real> O = ld<Real>(P4, N, c);
    const Pt<Real> iO = ld<Real>(P4, N, c - sx), jO = ld<Real>(P4, N, c - sy), kO = ld<Real>(P4, N, c - 1);
    Real m[9];
    Grp<Real> G;
    if (xf) {
        const Pt<Real> Oj = ld<Real>(P4, N, c + sy), iOj = ld<Real>(P4, N, c - sx + sy), ijO = ld<Real>(P4, N, c - sx - sy);
        const Pt<Real> Ok = ld<Real>(P4, N, c + 1), iOk = ld<Real>(P4, N, c - sx + 1), ikO = ld<Real>(P4, N, c - sx - 1);
        const D10<Real> X = d2(O, iO, invXI);
        const D10<Real> Y = d4(Oj, iOj, jO, ijO, inv4ET);
        const D10<Real> Z = d4(Ok, iOk, kO, ikO, inv4ZT);
#pragma unroll
        for (int q = 0; q < 9; q++) m[q] = (Real)mu.m[q][fid];
        groups(m, G);
        const Real Ux = Real(0.5) * (O.u + iO.u), Vx = Real(0.5) * (O.v + iO.v), Wx = Real(0.5) * (O.w + iO.w);
        const Real Tx = Real(0.5) * (O.T + iO.T);
        const Real mu_ = suth<Real>(p, Tx);
        const Real Jf = (Real)Ju[fid];
        FF[(0 * 3 + 0) * N + c] = mu_ * (G.a[1] * X.u + G.a[5] * X.v + G.a[7] * X.w + G.d[1] * Y.u + G.d[7] * Y.v +
                                         G.d[9] * Y.w + G.e[1] * Z.u + G.e[7] * Z.v + G.e[9] * Z.w) / Jf;
        FF[(1 * 3 + 0) * N + c] = mu_ * (G.a[5] * X.u + G.a[2] * X.v + G.a[6] * X.w + G.d[5] * Y.u + G.d[2] * Y.v +
                                         G.d[10] * Y.w + G.e[5] * Z.u + G.e[2] * Z.v + G.e[10] * Z.w) / Jf;
        FF[(2 * 3 + 0) * N + c] = mu_ * (G.a[7] * X.u + G.a[6] * X.v + G.a[3] * X.w + G.d[6] * Y.u + G.d[8] * Y.v +
                                         G.d[3] * Y.w + G.e[6] * Z.u + G.e[8] * Z.v + G.e[3] * Z.w) / Jf;
        FF[(3 * 3 + 0) * N + c] =
            mu_ * (Real(0.5) * G.a[1] * X.u2 + Real(0.5) * G.a[2] * X.v2 + Real(0.5) * G.a[3] * X.w2 + G.a[5] * X.uv + G.a[6] * X.vw +
                   G.a[7] * X.uw + CvK * G.a[4] * X.T / (Pr) + Real(0.5) * G.d[1] * Y.u2 + Real(0.5) * G.d[2] * Y.v2 +
                   Real(0.5) * G.d[3] * Y.w2 + G.d[5] * Vx * Y.u + G.d[6] * Wx * Y.u + G.d[7] * Ux * Y.v +
                   G.d[8] * Wx * Y.v + G.d[9] * Ux * Y.w + G.d[10] * Vx * Y.w + CvK * G.d[4] * Y.T / (Pr) +
                   Real(0.5) * G.e[1] * Z.u2 + Real(0.5) * G.e[2] * Z.v2 + Real(0.5) * G.e[3] * Z.w2 + G.e[5] * Vx * Z.u +
                   G.e[6] * Wx * Z.u + G.e[7] * Ux * Z.v + G.e[8] * Wx * Z.v + G.e[9] * Ux * Z.w +
                   G.e[10] * Vx * Z.w + CvK * G.e[4] * Z.T / (Pr)) /
            Jf;
    }
    if (yf) {
        const Pt<Real> Oi = ld<Real>(P4, N, c + sx), jOi = ld<Real>(P4, N, c + sx - sy), ijO = ld<Real>(P4, N, c - sx - sy);
        const Pt<Real> Ok = ld<Real>(P4, N, c + 1), jOk = ld<Real>(P4, N, c - sy + 1), jkO = ld<Real>(P4, N, c - sy - 1);
        const D10<Real> X = d4(Oi, jOi, iO, ijO, inv4XI);
        const D10<Real> Y = d2(O, jO, invET);
        const D10<Real> Z = d4(Ok, jOk, kO, jkO, inv4ZT);
#pragma unroll
        for (int q = 0; q < 9; q++) m[q] = (Real)mv.m[q][fid];
        groups(m, G);
        const Real Uy = Real(0.5) * (O.u + jO.u), Vy = Real(0.5) * (O.v + jO.v), Wy = Real(0.5) * (O.w + jO.w);
        const Real Ty = Real(0.5) * (O.T + jO.T);
        const Real mu_ = suth<Real>(p, Ty);
        const Real Jf = (Real)Jv[fid];
        FF[(0 * 3 + 1) * N + c] = mu_ * (G.d[1] * X.u + G.d[5] * X.v + G.d[6] * X.w + G.b[1] * Y.u + G.b[5] * Y.v +
                                         G.b[7] * Y.w + G.f[1] * Z.u + G.f[7] * Z.v + G.f[9] * Z.w) / Jf;
        FF[(1 * 3 + 1) * N + c] = mu_ * (G.d[7] * X.u + G.d[2] * X.v + G.d[8] * X.w + G.b[5] * Y.u + G.b[2] * Y.v +
                                         G.b[6] * Y.w + G.f[5] * Z.u + G.f[2] * Z.v + G.f[10] * Z.w) / Jf;
        FF[(2 * 3 + 1) * N + c] = mu_ * (G.d[9] * X.u + G.d[10] * X.v + G.d[3] * X.w + G.b[7] * Y.u + G.b[6] * Y.v +
                                         G.b[3] * Y.w + G.f[6] * Z.u + G.f[8] * Z.v + G.f[3] * Z.w) / Jf;
        FF[(3 * 3 + 1) * N + c] =
            mu_ * (Real(0.5) * G.b[1] * Y.u2 + Real(0.5) * G.b[2] * Y.v2 + Real(0.5) * G.b[3] * Y.w2 + G.b[5] * Y.uv + G.b[6] * Y.vw +
                   G.b[7] * Y.uw + CvK * G.b[4] * Y.T / (Pr) + Real(0.5) * G.d[1] * X.u2 + Real(0.5) * G.d[2] * X.v2 +
                   Real(0.5) * G.d[3] * X.w2 + G.d[5] * Uy * X.v + G.d[6] * Uy * X.w + G.d[7] * Vy * X.u +
                   G.d[8] * Vy * X.w + G.d[9] * Wy * X.u + G.d[10] * Wy * X.v + CvK * G.d[4] * X.T / (Pr) +
                   Real(0.5) * G.f[1] * Z.u2 + Real(0.5) * G.f[2] * Z.v2 + Real(0.5) * G.f[3] * Z.w2 + G.f[5] * Vy * Z.u +
                   G.f[6] * Wy * Z.u + G.f[7] * Uy * Z.v + G.f[8] * Wy * Z.v + G.f[9] * Uy * Z.w +
                   G.f[10] * Vy * Z.w + CvK * G.f[4] * Z.T / (Pr)) /
            Jf;
    }
    if (zf) {
        const Pt<Real> Oi = ld<Real>(P4, N, c + sx), kOi = ld<Real>(P4, N, c + sx - 1), ikO = ld<Real>(P4, N, c - sx - 1);
        const Pt<Real> Oj = ld<Real>(P4, N, c + sy), kOj = ld<Real>(P4, N, c + sy - 1), jkO = ld<Real>(P4, N, c - sy - 1);
        const Pt<Real> iOj = ld<Real>(P4, N, c - sx + sy);
        const D10<Real> X = d4(Oi, kOi, iO, ikO, inv4XI);
        D10<Real> Y = d4(Oj, kOj, jO, jkO, inv4ET);
        // Viscous_terms.cpp:744-746 pair kUj with iVj / iWj
        Y.uv = (Oj.u * Oj.v + kOj.u * iOj.v - jO.u * jO.v - jkO.u * jkO.v) * inv4ET;
        Y.vw = (Oj.v * Oj.w + kOj.v * iOj.w - jO.v * jO.w - jkO.v * jkO.w) * inv4ET;
        Y.uw = (Oj.u * Oj.w + kOj.u * iOj.w - jO.u * jO.w - jkO.u * jkO.w) * inv4ET;
        const D10<Real> Z = d2(O, kO, invZT);
#pragma unroll
        for (int q = 0; q < 9; q++) m[q] = (Real)mw.m[q][fid];
        groups(m, G);
        const Real Uz = Real(0.5) * (O.u + kO.u), Vz = Real(0.5) * (O.v + kO.v), Wz = Real(0.5) * (O.w + kO.w);
        const Real Tz = Real(0.5) * (O.T + kO.T);
        const Real mu_ = suth<Real>(p, Tz);
        const Real Jf = (Real)Jw[fid];
        FF[(0 * 3 + 2) * N + c] = mu_ * (G.e[1] * X.u + G.e[5] * X.v + G.e[6] * X.w + G.f[1] * Y.u + G.f[5] * Y.v +
                                         G.f[6] * Y.w + G.c[1] * Z.u + G.c[5] * Z.v + G.c[7] * Z.w) / Jf;
        FF[(1 * 3 + 2) * N + c] = mu_ * (G.e[7] * X.u + G.e[2] * X.v + G.e[8] * X.w + G.f[7] * Y.u + G.f[2] * Y.v +
                                         G.f[8] * Y.w + G.c[5] * Z.u + G.c[2] * Z.v + G.c[6] * Z.w) / Jf;
        FF[(2 * 3 + 2) * N + c] = mu_ * (G.e[9] * X.u + G.e[10] * X.v + G.e[3] * X.w + G.f[9] * Y.u + G.f[10] * Y.v +
                                         G.f[3] * Y.w + G.c[7] * Z.u + G.c[6] * Z.v + G.c[3] * Z.w) / Jf;
        FF[(3 * 3 + 2) * N + c] =
            mu_ * (Real(0.5) * G.c[1] * Z.u2 + Real(0.5) * G.c[2] * Z.v2 + Real(0.5) * G.c[3] * Z.w2 + G.c[5] * Z.uv + G.c[6] * Z.vw +
                   G.c[7] * Z.uw + CvK * G.c[4] * Z.T / (Pr) + Real(0.5) * G.e[1] * X.u2 + Real(0.5) * G.e[2] * X.v2 +
                   Real(0.5) * G.e[3] * X.w2 + G.e[5] * Uz * X.v + G.e[6] * Uz * X.w + G.e[7] * Vz * X.u +
                   G.e[8] * Vz * X.w + G.e[9] * Wz * X.u + G.e[10] * Wz * X.v + CvK * G.e[4] * X.T / (Pr) +
                   Real(0.5) * G.f[1] * Y.u2 + Real(0.5) * G.f[2] * Y.v2 + Real(0.5) * G.f[3] * Y.w2 + G.f[5] * Uz * Y.v +
                   G.f[6] * Uz * Y.w + G.f[7] * Vz * Y.u + G.f[8] * Vz * Y.w + G.f[9] * Wz * Y.u +
                   G.f[10] * Wz * Y.v + CvK * G.f[4] * Y.T / (Pr)) /
            Jf;
    }
    (void)ix;
}

__global__ void __launch_bounds__(256) k_visc_div(Geo g, Phys p, const double *__restrict__ FF, F4 vF)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long N = g.N, sx = g.Ym * g.Zm, sy = g.Zm;
    const long c = ix(g, i, j, k);
    const double invXI = 1. / (p.dxi), invET = 1. / (p.deta), invZT = 1. / (p.dzeta);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double *X = FF + (v * 3 + 0) * N, *Y = FF + (v * 3 + 1) * N, *Z = FF + (v * 3 + 2) * N;
        vF.v[v][c] = (X[c + sx] - X[c]) * invXI + (Y[c + sy] - Y[c]) * invET + (Z[c + 1] - Z[c]) * invZT;
    }
}

void launch_viscous(const Geo &g, const Phys &p, C5 U, const double *J, const double *Ju, M9 mu,
                    const double *Jv, M9 mv, const double *Jw, M9 mw, double *prim4, double *faces12,
                    F4 vF, hipStream_t s)
{
    const dim3 b1(256);
    const int np = g.NZZ * g.NYY, nf = (g.NZZ - 1) * (g.NYY - 1);
    dim3 gp((np + 255) / 256, 1, g.gend);   // i=2..gend+1
    hipLaunchKernelGGL(k_prim4, gp, b1, 0, s, g, p.K, p.R, U, J, prim4);
    dim3 gf((nf + 255) / 256, 1, g.gend - 1);   // i=3..gend+1
    if (p.fp32) {
        hipLaunchKernelGGL((k_visc_faces<float, 0>), gf, b1, 0, s, g, p, prim4, Ju, mu, Jv, mv, Jw, mw, faces12);
        hipLaunchKernelGGL((k_visc_faces<float, 1>), gf, b1, 0, s, g, p, prim4, Ju, mu, Jv, mv, Jw, mw, faces12);
        hipLaunchKernelGGL((k_visc_faces<float, 2>), gf, b1, 0, s, g, p, prim4, Ju, mu, Jv, mv, Jw, mw, faces12);
    } else {
        hipLaunchKernelGGL((k_visc_faces<double, 0>), gf, b1, 0, s, g, p, prim4, Ju, mu, Jv, mv, Jw, mw, faces12);
        hipLaunchKernelGGL((k_visc_faces<double, 1>), gf, b1, 0, s, g, p, prim4, Ju, mu, Jv, mv, Jw, mw, faces12);
        hipLaunchKernelGGL((k_visc_faces<double, 2>), gf, b1, 0, s, g, p, prim4, Ju, mu, Jv, mv, Jw, mw, faces12);
    }
    dim3 b(64, 4);
    dim3 gd((g.NZ - 1 + b.x - 1) / b.x, (g.NY - 1 + b.y - 1) / b.y, g.gend - 2);   // i=3..gend
    hipLaunchKernelGGL(k_visc_div, gd, b, 0, s, g, p, faces12, vF);
}

}  // namespace cgli
