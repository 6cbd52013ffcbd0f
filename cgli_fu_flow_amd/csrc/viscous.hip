// viscous.hip -- viscous and heat-flux terms on gfx950, templated on the
// arithmetic type of the face fluxes (double: the reference path; float: the
// FP32 residual path, reading the FP64 primitives and metrics and summing
// the divergence of FP64-stored face fluxes).
//
// The reference's three stages (Viscous_terms.cpp:183-905) in one kernel:
//   1. primitives (u, v, w, T) of every cell the stencil touches, once per
//      block;
//   2. per cell, the fluxes through its low x, y and z faces, from
//      second-order face gradients with four-point cross averages and
//      Sutherland viscosity;
//   3. vF = dF/dxi + dG/deta + dH/dzeta on interior cells.
// A block owns a (j, k) tile and marches along i through a segment of
// planes: a ring of four planes of primitives (tile plus a 1-cell halo) and
// the tile's y and z face fluxes of two planes live in LDS, the x face flux
// of the cell below travels in registers, so that U and each face's J and
// metrics are read once and vF is the only store.  A plane's divergence is
// taken after the next plane's barrier, so one barrier per plane orders
// every LDS access.  Loads run ahead of their use: a
// plane's conserved variables one plane ahead, each face's metrics one face
// ahead, so that the march seldom waits on memory.  The arithmetic is the
// reference's except for divisions: a face's 1/Jf and the constant Cv K / Pr
// are formed once and multiplied, the Sutherland power (T/Ts)^1.5 is
// x sqrt(x), and the temperature is formed without the J that cancels from
// it; results agree with the reference's to rounding.
#include <algorithm>

#include "cgli_internal.hpp"

// Products and sums in this file contract to fused multiply-adds (the
// face fluxes are long dot products of metric groups and gradients: 7 %
// less time at 256x192x256); the results agree with the separately rounded
// reference to rounding.
#pragma clang fp contract(fast)

namespace cgli {

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
    // (T / Ts)^1.5 as x sqrt(x): within an ulp or two of pow
    const Real Ts = (Real)p.suth_T, S = (Real)p.suth_S;
    const Real x = T / Ts;
    return (Real)p.mu_L * (x * sqrt(x)) * (Ts + S) / (T + S);
}

// The face fluxes (variables 2..5 of the reference, here 0..3).  x face
// i-1/2 of cell (i, j, k): O = P(i, j, k), iO = P(i-1, j, k), and the j and k
// neighbours of both; Jf and m are the face point's Jacobian and metrics.
template <typename Real>
__device__ __forceinline__ void xflux(const Phys &p, const Pt<Real> &O, const Pt<Real> &iO, const Pt<Real> &Oj,
                                      const Pt<Real> &iOj, const Pt<Real> &jO, const Pt<Real> &ijO, const Pt<Real> &Ok,
                                      const Pt<Real> &iOk, const Pt<Real> &kO, const Pt<Real> &ikO, const Real m[9],
                                      Real Jf, double F[4])
{
    const Real invXI = (Real)(1. / (p.dxi)), inv4ET = (Real)(1. / (4 * p.deta)), inv4ZT = (Real)(1. / (4 * p.dzeta));
    const Real CvKPr = (Real)(p.Cv * p.K / p.Pr_L), iJf = Real(1) / Jf;
    Grp<Real> G;
    const D10<Real> X = d2(O, iO, invXI);
    const D10<Real> Y = d4(Oj, iOj, jO, ijO, inv4ET);
    const D10<Real> Z = d4(Ok, iOk, kO, ikO, inv4ZT);
    groups(m, G);
    const Real Ux = Real(0.5) * (O.u + iO.u), Vx = Real(0.5) * (O.v + iO.v), Wx = Real(0.5) * (O.w + iO.w);
    const Real Tx = Real(0.5) * (O.T + iO.T);
    const Real mu_ = suth<Real>(p, Tx);
    F[0] = mu_ * (G.a[1] * X.u + G.a[5] * X.v + G.a[7] * X.w + G.d[1] * Y.u + G.d[7] * Y.v + G.d[9] * Y.w +
                  G.e[1] * Z.u + G.e[7] * Z.v + G.e[9] * Z.w) * iJf;
    F[1] = mu_ * (G.a[5] * X.u + G.a[2] * X.v + G.a[6] * X.w + G.d[5] * Y.u + G.d[2] * Y.v + G.d[10] * Y.w +
                  G.e[5] * Z.u + G.e[2] * Z.v + G.e[10] * Z.w) * iJf;
    F[2] = mu_ * (G.a[7] * X.u + G.a[6] * X.v + G.a[3] * X.w + G.d[6] * Y.u + G.d[8] * Y.v + G.d[3] * Y.w +
                  G.e[6] * Z.u + G.e[8] * Z.v + G.e[3] * Z.w) * iJf;
    F[3] = mu_ * (Real(0.5) * G.a[1] * X.u2 + Real(0.5) * G.a[2] * X.v2 + Real(0.5) * G.a[3] * X.w2 + G.a[5] * X.uv +
                  G.a[6] * X.vw + G.a[7] * X.uw + CvKPr * G.a[4] * X.T + Real(0.5) * G.d[1] * Y.u2 +
                  Real(0.5) * G.d[2] * Y.v2 + Real(0.5) * G.d[3] * Y.w2 + G.d[5] * Vx * Y.u + G.d[6] * Wx * Y.u +
                  G.d[7] * Ux * Y.v + G.d[8] * Wx * Y.v + G.d[9] * Ux * Y.w + G.d[10] * Vx * Y.w +
                  CvKPr * G.d[4] * Y.T + Real(0.5) * G.e[1] * Z.u2 + Real(0.5) * G.e[2] * Z.v2 +
                  Real(0.5) * G.e[3] * Z.w2 + G.e[5] * Vx * Z.u + G.e[6] * Wx * Z.u + G.e[7] * Ux * Z.v +
                  G.e[8] * Wx * Z.v + G.e[9] * Ux * Z.w + G.e[10] * Vx * Z.w + CvKPr * G.e[4] * Z.T) *
           iJf;
}

// y face j-1/2 of cell (i, j, k): O = P(i, j, k), jO = P(i, j-1, k), and the
// i and k neighbours of both
template <typename Real>
__device__ __forceinline__ void yflux(const Phys &p, const Pt<Real> &O, const Pt<Real> &jO, const Pt<Real> &Oi,
                                      const Pt<Real> &jOi, const Pt<Real> &iO, const Pt<Real> &ijO, const Pt<Real> &Ok,
                                      const Pt<Real> &jOk, const Pt<Real> &kO, const Pt<Real> &jkO, const Real m[9],
                                      Real Jf, double F[4])
{
    const Real inv4XI = (Real)(1. / (4 * p.dxi)), invET = (Real)(1. / (p.deta)), inv4ZT = (Real)(1. / (4 * p.dzeta));
    const Real CvKPr = (Real)(p.Cv * p.K / p.Pr_L), iJf = Real(1) / Jf;
    Grp<Real> G;
    const D10<Real> X = d4(Oi, jOi, iO, ijO, inv4XI);
    const D10<Real> Y = d2(O, jO, invET);
    const D10<Real> Z = d4(Ok, jOk, kO, jkO, inv4ZT);
    groups(m, G);
    const Real Uy = Real(0.5) * (O.u + jO.u), Vy = Real(0.5) * (O.v + jO.v), Wy = Real(0.5) * (O.w + jO.w);
    const Real Ty = Real(0.5) * (O.T + jO.T);
    const Real mu_ = suth<Real>(p, Ty);
    F[0] = mu_ * (G.d[1] * X.u + G.d[5] * X.v + G.d[6] * X.w + G.b[1] * Y.u + G.b[5] * Y.v + G.b[7] * Y.w +
                  G.f[1] * Z.u + G.f[7] * Z.v + G.f[9] * Z.w) * iJf;
    F[1] = mu_ * (G.d[7] * X.u + G.d[2] * X.v + G.d[8] * X.w + G.b[5] * Y.u + G.b[2] * Y.v + G.b[6] * Y.w +
                  G.f[5] * Z.u + G.f[2] * Z.v + G.f[10] * Z.w) * iJf;
    F[2] = mu_ * (G.d[9] * X.u + G.d[10] * X.v + G.d[3] * X.w + G.b[7] * Y.u + G.b[6] * Y.v + G.b[3] * Y.w +
                  G.f[6] * Z.u + G.f[8] * Z.v + G.f[3] * Z.w) * iJf;
    F[3] = mu_ * (Real(0.5) * G.b[1] * Y.u2 + Real(0.5) * G.b[2] * Y.v2 + Real(0.5) * G.b[3] * Y.w2 + G.b[5] * Y.uv +
                  G.b[6] * Y.vw + G.b[7] * Y.uw + CvKPr * G.b[4] * Y.T + Real(0.5) * G.d[1] * X.u2 +
                  Real(0.5) * G.d[2] * X.v2 + Real(0.5) * G.d[3] * X.w2 + G.d[5] * Uy * X.v + G.d[6] * Uy * X.w +
                  G.d[7] * Vy * X.u + G.d[8] * Vy * X.w + G.d[9] * Wy * X.u + G.d[10] * Wy * X.v +
                  CvKPr * G.d[4] * X.T + Real(0.5) * G.f[1] * Z.u2 + Real(0.5) * G.f[2] * Z.v2 +
                  Real(0.5) * G.f[3] * Z.w2 + G.f[5] * Vy * Z.u + G.f[6] * Wy * Z.u + G.f[7] * Uy * Z.v +
                  G.f[8] * Wy * Z.v + G.f[9] * Uy * Z.w + G.f[10] * Vy * Z.w + CvKPr * G.f[4] * Z.T) *
           iJf;
}

// z face k-1/2 of cell (i, j, k): O = P(i, j, k), kO = P(i, j, k-1), and the
// i and j neighbours of both; iOj = P(i-1, j+1, k) enters through the
// reference's pairing of kUj with iVj / iWj (Viscous_terms.cpp:744-746)
template <typename Real>
__device__ __forceinline__ void zflux(const Phys &p, const Pt<Real> &O, const Pt<Real> &kO, const Pt<Real> &Oi,
                                      const Pt<Real> &kOi, const Pt<Real> &iO, const Pt<Real> &ikO, const Pt<Real> &Oj,
                                      const Pt<Real> &kOj, const Pt<Real> &jO, const Pt<Real> &jkO, const Pt<Real> &iOj,
                                      const Real m[9], Real Jf, double F[4])
{
    const Real inv4XI = (Real)(1. / (4 * p.dxi)), inv4ET = (Real)(1. / (4 * p.deta)), invZT = (Real)(1. / (p.dzeta));
    const Real CvKPr = (Real)(p.Cv * p.K / p.Pr_L), iJf = Real(1) / Jf;
    Grp<Real> G;
    const D10<Real> X = d4(Oi, kOi, iO, ikO, inv4XI);
    D10<Real> Y = d4(Oj, kOj, jO, jkO, inv4ET);
    Y.uv = (Oj.u * Oj.v + kOj.u * iOj.v - jO.u * jO.v - jkO.u * jkO.v) * inv4ET;
    Y.vw = (Oj.v * Oj.w + kOj.v * iOj.w - jO.v * jO.w - jkO.v * jkO.w) * inv4ET;
    Y.uw = (Oj.u * Oj.w + kOj.u * iOj.w - jO.u * jO.w - jkO.u * jkO.w) * inv4ET;
    const D10<Real> Z = d2(O, kO, invZT);
    groups(m, G);
    const Real Uz = Real(0.5) * (O.u + kO.u), Vz = Real(0.5) * (O.v + kO.v), Wz = Real(0.5) * (O.w + kO.w);
    const Real Tz = Real(0.5) * (O.T + kO.T);
    const Real mu_ = suth<Real>(p, Tz);
    F[0] = mu_ * (G.e[1] * X.u + G.e[5] * X.v + G.e[6] * X.w + G.f[1] * Y.u + G.f[5] * Y.v + G.f[6] * Y.w +
                  G.c[1] * Z.u + G.c[5] * Z.v + G.c[7] * Z.w) * iJf;
    F[1] = mu_ * (G.e[7] * X.u + G.e[2] * X.v + G.e[8] * X.w + G.f[7] * Y.u + G.f[2] * Y.v + G.f[8] * Y.w +
                  G.c[5] * Z.u + G.c[2] * Z.v + G.c[6] * Z.w) * iJf;
    F[2] = mu_ * (G.e[9] * X.u + G.e[10] * X.v + G.e[3] * X.w + G.f[9] * Y.u + G.f[10] * Y.v + G.f[3] * Y.w +
                  G.c[7] * Z.u + G.c[6] * Z.v + G.c[3] * Z.w) * iJf;
    F[3] = mu_ * (Real(0.5) * G.c[1] * Z.u2 + Real(0.5) * G.c[2] * Z.v2 + Real(0.5) * G.c[3] * Z.w2 + G.c[5] * Z.uv +
                  G.c[6] * Z.vw + G.c[7] * Z.uw + CvKPr * G.c[4] * Z.T + Real(0.5) * G.e[1] * X.u2 +
                  Real(0.5) * G.e[2] * X.v2 + Real(0.5) * G.e[3] * X.w2 + G.e[5] * Uz * X.v + G.e[6] * Uz * X.w +
                  G.e[7] * Vz * X.u + G.e[8] * Vz * X.w + G.e[9] * Wz * X.u + G.e[10] * Wz * X.v +
                  CvKPr * G.e[4] * X.T + Real(0.5) * G.f[1] * Y.u2 + Real(0.5) * G.f[2] * Y.v2 +
                  Real(0.5) * G.f[3] * Y.w2 + G.f[5] * Uz * Y.v + G.f[6] * Uz * Y.w + G.f[7] * Vz * Y.u +
                  G.f[8] * Vz * Y.w + G.f[9] * Wz * Y.u + G.f[10] * Wz * Y.v + CvKPr * G.f[4] * Y.T) *
           iJf;
}

// Tile of the fused kernel: the threads are VT_J x VT_K face positions (j, k)
// (k fastest); the cells whose four side faces all lie among them, the
// first VT_J - 1 rows and VT_K - 1 columns, are the tile's output.
constexpr int VT_J = 8, VT_K = 32, VT_THREADS = VT_J * VT_K;
constexpr int VH_J = VT_J + 2, VH_K = VT_K + 2;   // staged primitives: positions plus a 1-cell halo

// The conserved variables of a plane's halo positions, two per thread,
// loaded a plane ahead of their conversion.  The loads run for every slot
// (a slot past the halo, or a position outside the primitive box, reads the
// nearest position inside it and is not stored): no branch around them, so
// that the compiler's wait counts stay exact.
static_assert(VH_J * VH_K <= 2 * VT_THREADS, "two halo positions per thread");
struct Uplane {
    double r[2][5];
};

__device__ __forceinline__ void u_load(const Geo &g, C5 U, int i, int j0, int k0, Uplane &R)
{
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int t = min((int)threadIdx.x + s * VT_THREADS, VH_J * VH_K - 1);
        const int j = min(j0 - 1 + t / VH_K, g.NYY), k = min(k0 - 1 + t % VH_K, g.NZZ);
        const long c = ix(g, i, j, k);
#pragma unroll
        for (int v = 0; v < 5; v++) R.r[s][v] = U.v[v][c];
    }
}

// primitives (u, v, w, T) of plane i of the tile's halo, from its
// conserved variables R, into ring slot i & 3.
// The reference's T = P / (rho R) / J with P = (U5 - rho VV / 2)(K - 1) J
// (Viscous_terms.cpp:260-266); J cancels, so T is formed from the stored
// state alone, with one division per cell.
__device__ __forceinline__ void visc_plane(const Geo &g, const Phys &p, const Uplane &R, int i, int j0, int k0,
                                           double (*sP)[4][VH_J][VH_K])
{
    const double KR = (p.K - 1) / p.R;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int t = threadIdx.x + s * VT_THREADS;
        const int hj = t / VH_K, hk = t % VH_K;
        // outside the primitive box: never read for a used face
        if (t >= VH_J * VH_K || j0 - 1 + hj > g.NYY || k0 - 1 + hk > g.NZZ) continue;
        const double r = R.r[s][0], ir = 1.0 / r;
        const double u = R.r[s][1] * ir, v = R.r[s][2] * ir, w = R.r[s][3] * ir;
        const double T = (R.r[s][4] - 0.5 * r * (u * u + v * v + w * w)) * KR * ir;
        double(*pl)[VH_J][VH_K] = sP[i & 3];
        pl[0][hj][hk] = u;
        pl[1][hj][hk] = v;
        pl[2][hj][hk] = w;
        pl[3][hj][hk] = T;
    }
}

template <typename Real>
__device__ __forceinline__ Pt<Real> lp(const double (*sP)[4][VH_J][VH_K], int i, int hj, int hk)
{
    const double(*pl)[VH_J][VH_K] = sP[i & 3];
    return Pt<Real>{(Real)pl[0][hj][hk], (Real)pl[1][hj][hk], (Real)pl[2][hj][hk], (Real)pl[3][hj][hk]};
}

// a face's metrics and Jacobian
template <typename Real>
struct Mf {
    Real m[9], J;
};

template <typename Real>
__device__ __forceinline__ void ldm(Mf<Real> &M, const M9 &mm, const double *__restrict__ Jx, long fid)
{
#pragma unroll
    for (int q = 0; q < 9; q++) M.m[q] = (Real)mm.m[q][fid];
    M.J = (Real)Jx[fid];
}

// x face of cell (i, j, k) (between planes i-1 and i) at the thread's position
template <typename Real>
__device__ __forceinline__ void xface(const Phys &p, const double (*sP)[4][VH_J][VH_K], int i, int hj, int hk,
                                      const Mf<Real> &M, double F[4])
{
    xflux<Real>(p, lp<Real>(sP, i, hj, hk), lp<Real>(sP, i - 1, hj, hk), lp<Real>(sP, i, hj + 1, hk),
                lp<Real>(sP, i - 1, hj + 1, hk), lp<Real>(sP, i, hj - 1, hk), lp<Real>(sP, i - 1, hj - 1, hk),
                lp<Real>(sP, i, hj, hk + 1), lp<Real>(sP, i - 1, hj, hk + 1), lp<Real>(sP, i, hj, hk - 1),
                lp<Real>(sP, i - 1, hj, hk - 1), M.m, M.J, F);
}

// vF of the interior cells.  Grid: (tiles in k, tiles in j, segments of seg
// planes in i); every face the reference's divergence reads is computed
// (x faces i <= gend+1, y faces j <= NYY, z faces k <= NZZ).
template <typename Real>
__global__ void __launch_bounds__(VT_THREADS, 2) k_viscous(Geo g, Phys p, C5 U, const double *__restrict__ Ju, M9 mu,
                                                        const double *__restrict__ Jv, M9 mv,
                                                        const double *__restrict__ Jw, M9 mw, F4 vF, int seg)
{
    // primitives of planes i-1 .. i+2 (slot i & 3): the ring's fourth slot
    // and the two parities of the face fluxes let one barrier per plane
    // order every LDS access
    __shared__ double sP[4][4][VH_J][VH_K];
    __shared__ double sF[2][2][4][VT_J][VT_K];   // y and z face fluxes of planes of either parity
    const int tk = threadIdx.x % VT_K, tj = threadIdx.x / VT_K;
    const int j0 = 2 + blockIdx.y * (VT_J - 1), k0 = 2 + blockIdx.x * (VT_K - 1);
    const int j = j0 + tj, k = k0 + tk, hj = tj + 1, hk = tk + 1;
    const int iS = 3 + blockIdx.z * seg, iE = min(iS + seg, g.gend + 1);
    const bool out = tj < VT_J - 1 && tk < VT_K - 1 && j <= g.NY && k <= g.NZ;   // an output cell
    const bool yok = j <= g.NYY && k <= g.NZ, zok = j <= g.NY && k <= g.NZZ;   // its low y / z face exists
    const double invXI = 1. / (p.dxi), invET = 1. / (p.deta), invZT = 1. / (p.dzeta);
    Uplane R;
    u_load(g, U, iS - 1, j0, k0, R);
    visc_plane(g, p, R, iS - 1, j0, k0, sP);
    u_load(g, U, iS, j0, k0, R);
    visc_plane(g, p, R, iS, j0, k0, sP);
    u_load(g, U, iS + 1, j0, k0, R);
    // face metrics are loaded one face ahead of their use: the x face of
    // the next plane while the z face computes, the y face while the x face
    // computes, the z face while the y face computes
    Mf<Real> MX, MY, MZ;
    // the loads run for every thread (no branch around them, so that the
    // compiler's wait counts stay exact); a position without the face reads
    // the nearest one that has it and ignores the values
    const long cjk = (long)(min(j, g.NYY) - 1) * g.Zm + (min(k, g.NZZ) - 1), sx = (long)g.Ym * g.Zm;
    ldm<Real>(MX, mu, Ju, (iS - 1) * sx + cjk);
    __syncthreads();
    double Fx[4] = {0.0, 0.0, 0.0, 0.0};   // x face below the current cell
    double dX[4] = {0.0, 0.0, 0.0, 0.0};   // the x part of the previous plane's divergence
    if (out) xface<Real>(p, sP, iS, hj, hk, MX, Fx);
    ldm<Real>(MX, mu, Ju, iS * sx + cjk);
    // vF of plane q from its faces (the y and z fluxes in sF[q & 1])
    auto divergence = [&](int q) {
        if (!out) return;
        const long c = ix(g, q, j, k);
        const double(*fy)[VT_J][VT_K] = sF[q & 1][0];
        const double(*fz)[VT_J][VT_K] = sF[q & 1][1];
#pragma unroll
        for (int v = 0; v < 4; v++)
            vF.v[v][c] = dX[v] + (fy[v][tj + 1][tk] - fy[v][tj][tk]) * invET + (fz[v][tj][tk + 1] - fz[v][tj][tk]) * invZT;
    };
    for (int i = iS; i < iE; i++) {
        const long fid = (i - 1) * sx + cjk;
        visc_plane(g, p, R, i + 1, j0, k0, sP);
        ldm<Real>(MY, mv, Jv, fid);
        __syncthreads();
        if (i > iS) divergence(i - 1);
        // the next plane's conserved variables, converted at the top of the
        // next iteration (the last iteration reloads plane iE, unused)
        u_load(g, U, min(i + 2, iE), j0, k0, R);
        // the three faces one after the other: the scheduling barriers keep the
        // compiler from hoisting all their stencil reads at once
        double Fh[4] = {0.0, 0.0, 0.0, 0.0};
        if (out) xface<Real>(p, sP, i + 1, hj, hk, MX, Fh);
        ldm<Real>(MZ, mw, Jw, fid);
        __builtin_amdgcn_sched_barrier(0);
        if (yok) {
            double F[4];
            yflux<Real>(p, lp<Real>(sP, i, hj, hk), lp<Real>(sP, i, hj - 1, hk), lp<Real>(sP, i + 1, hj, hk),
                        lp<Real>(sP, i + 1, hj - 1, hk), lp<Real>(sP, i - 1, hj, hk), lp<Real>(sP, i - 1, hj - 1, hk),
                        lp<Real>(sP, i, hj, hk + 1), lp<Real>(sP, i, hj - 1, hk + 1), lp<Real>(sP, i, hj, hk - 1),
                        lp<Real>(sP, i, hj - 1, hk - 1), MY.m, MY.J, F);
#pragma unroll
            for (int v = 0; v < 4; v++) sF[i & 1][0][v][tj][tk] = F[v];
        }
        ldm<Real>(MX, mu, Ju, (min(i + 1, iE - 1)) * sx + cjk);
        __builtin_amdgcn_sched_barrier(0);
        if (zok) {
            double F[4];
            zflux<Real>(p, lp<Real>(sP, i, hj, hk), lp<Real>(sP, i, hj, hk - 1), lp<Real>(sP, i + 1, hj, hk),
                        lp<Real>(sP, i + 1, hj, hk - 1), lp<Real>(sP, i - 1, hj, hk), lp<Real>(sP, i - 1, hj, hk - 1),
                        lp<Real>(sP, i, hj + 1, hk), lp<Real>(sP, i, hj + 1, hk - 1), lp<Real>(sP, i, hj - 1, hk),
                        lp<Real>(sP, i, hj - 1, hk - 1), lp<Real>(sP, i - 1, hj + 1, hk), MZ.m, MZ.J, F);
#pragma unroll
            for (int v = 0; v < 4; v++) sF[i & 1][1][v][tj][tk] = F[v];
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            dX[v] = (Fh[v] - Fx[v]) * invXI;
            Fx[v] = Fh[v];
        }
    }
    __syncthreads();
    divergence(iE - 1);
}

void launch_viscous(const Geo &g, const Phys &p, C5 U, const double *J, const double *Ju, M9 mu,
                    const double *Jv, M9 mv, const double *Jw, M9 mw, F4 vF, hipStream_t s)
{
    const int tj = (g.NY - 1 + VT_J - 2) / (VT_J - 1), tk = (g.NZ - 1 + VT_K - 2) / (VT_K - 1);
    const int ni = g.gend - 2;   // cell planes i = 3..gend
    // segments of at least 8 planes, enough blocks to fill the chip three times over
    int seg = (int)std::min<long>(ni, std::max<long>(8, (long)ni * tj * tk / 2304));
    const dim3 gr(tk, tj, (ni + seg - 1) / seg), bl(VT_THREADS);
    if (p.fp32)
        hipLaunchKernelGGL(k_viscous<float>, gr, bl, 0, s, g, p, U, Ju, mu, Jv, mv, Jw, mw, vF, seg);
    else
        hipLaunchKernelGGL(k_viscous<double>, gr, bl, 0, s, g, p, U, Ju, mu, Jv, mv, Jw, mw, vF, seg);
}

}  // namespace cgli
