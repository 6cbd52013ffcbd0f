// flux.hip -- inviscid face fluxes on gfx950.
//
// Templated on the arithmetic type T: double (the reference path) or float
// (the FP32 residual path: fields stay FP64 in HBM, each thread converts its
// stencil on load and the face flux back on store).
// Each face's J-weighted MUSCL-5 left and right states are rebuilt from the
// six cells along the sweep direction (Flux_X.cpp:106-216,
// Flux_Y.cpp:106-249, Flux_Z.cpp:105-216), and the low-Mach preconditioned
// Roe flux (Flux_X.cpp:229-350) is taken between them.  The reference's
// ML/MR scratch never goes to HBM.  x and y faces: a thread marches along
// the sweep direction through a segment of faces, with the six cells of the
// current face (conserved variables and 1/J) in registers and the next cell
// and face metrics in flight, so each cell is read once per segment.  z
// faces (the direction of contiguous memory): one thread per face, which
// issues all of its face's loads before it waits for the first.  The
// expressions are the reference's, except that a divisor used more than once
// is inverted once and multiplied (the FP64 division is the costliest
// instruction sequence of the kernel) and products and sums fuse: results
// agree with the CPU build to rounding (a few ulps), not bit for bit.
// Optional dissipation scaling (the older flux files, Flux_X~~.cpp:552-556):
// each face's Roe dissipation Fav is multiplied by epS, or by EpF[face] (the
// filter sensor, face-indexed like the fluxes) when EpF is given.  The live
// reference has epS = 1 and no EpF; multiplying by 1 is exact.
#include <algorithm>

#include "cgli_internal.hpp"

// Products and sums in this file contract to fused multiply-adds; the
// results agree with the separately rounded reference to rounding.
#pragma clang fp contract(fast)

namespace cgli {

// J-weighted MUSCL reconstruction of one face's left and right states from
// the six cells c0 + (m - 2) s, m = 0..5, along the sweep direction.  Each
// stencil shape of the reference (5-point, 3rd order, two-cell average;
// Flux_X.cpp:106-216) is its coefficient vector over the six positions:
// the state is sum(c[m] U[m] / ...) weighted by temp = 1 / sum(c[m] / J[m])
// (Flux_X.cpp:118-121).  A position neither stencil uses reads cell c0
// instead (its coefficient is zero; the substitute keeps the load in bounds
// and finite), so every load of the face is issued before the first use.
template <typename T>
__device__ __forceinline__ T ld(const double *p, long i)
{
    return (T)p[i];
}

template <typename T>
__device__ __forceinline__ void shape(T c[6], T c0, T c1, T c2, T c3, T c4, T c5)
{
    c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3; c[4] = c4; c[5] = c5;
}
// the shapes by cell position
template <typename T> __device__ __forceinline__ void sh5L(T c[6]) { shape<T>(c, 2, -13, 47, 27, -3, 0); }
template <typename T> __device__ __forceinline__ void sh5R(T c[6]) { shape<T>(c, 0, -3, 27, 47, -13, 2); }
// 3rd-order left state from cells 1, 2, 3 and right state from 2, 3, 4
template <typename T> __device__ __forceinline__ void sh3L(T c[6]) { shape<T>(c, 0, -1, 5, 2, 0, 0); }
template <typename T> __device__ __forceinline__ void sh3R(T c[6]) { shape<T>(c, 0, 0, 2, 5, -1, 0); }
// two-cell averages
template <typename T> __device__ __forceinline__ void sh12(T c[6]) { shape<T>(c, 0, 1, 1, 0, 0, 0); }
template <typename T> __device__ __forceinline__ void sh23(T c[6]) { shape<T>(c, 0, 0, 1, 1, 0, 0); }
template <typename T> __device__ __forceinline__ void sh34(T c[6]) { shape<T>(c, 0, 0, 0, 1, 1, 0); }

template <typename T>
__device__ __forceinline__ T dot6(const T c[6], const T x[6])
{
    T s = c[0] * x[0];
#pragma unroll
    for (int m = 1; m < 6; m++) s += c[m] * x[m];
    return s;
}

// the six cells' J and conserved variables of one face, loaded together
template <typename T>
struct Sten {
    T q[6], u[5][6];
};

template <typename T>
__device__ __forceinline__ void load_sten(C5 U, const double *__restrict__ J, long c0, long s, const T cl[6],
                                          const T cr[6], Sten<T> &S)
{
    long at[6];
#pragma unroll
    for (int m = 0; m < 6; m++) at[m] = (cl[m] != T(0) || cr[m] != T(0)) ? c0 + (m - 2) * s : c0;
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
        for (int m = 0; m < 6; m++) S.u[v][m] = ld<T>(U.v[v], at[m]);
#pragma unroll
    for (int m = 0; m < 6; m++) S.q[m] = ld<T>(J, at[m]);
}

template <typename T>
__device__ __forceinline__ void face_states(Sten<T> &S, const T cl[6], const T cr[6], T Lq[5], T Rq[5])
{
#pragma unroll
    for (int m = 0; m < 6; m++) S.q[m] = T(1) / S.q[m];
    const T wl = T(1) / dot6(cl, S.q), wr = T(1) / dot6(cr, S.q);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        Lq[v] = wl * dot6(cl, S.u[v]);
        Rq[v] = wr * dot6(cr, S.u[v]);
    }
}

// Preconditioned Roe flux through one face (Flux_X.cpp:251-350).  (mx,my,mz)
// are the face metrics, (nx_,ny_,nz_) the vector multiplying the pressure
// jump in the momentum dissipation.  The reference's expressions, with each
// divisor that recurs (the densities, the Roe weight sum, K - 1, C, S, Jf)
// inverted once and multiplied, and the dissipation's division by
// 1/sqrt(|m|^2) taken as a product with sqrt(|m|^2): the results differ from
// the reference's by rounding only.
template <typename T>
__device__ __forceinline__ void roe_face(const T L[5], const T Rq[5], T mx, T my, T mz, T Jf, T e, T K, T nx_,
                                         T ny_, T nz_, T ep, T F[5], T &pL_out, T &pR_out, T &Fav3_out)
{
    const T sq = sqrt(mx * mx + my * my + mz * mz);
    const T insqr = T(1.0) / sq;
    const T Km1 = K - T(1), iKm1 = T(1) / Km1;

    const T rL = L[0], irL = T(1) / rL;
    const T uL = L[1] * irL, vL = L[2] * irL, wL = L[3] * irL;
    const T UnL = mx * uL + my * vL + mz * wL;
    const T qL = uL * uL + vL * vL + wL * wL;
    const T pL = (L[4] - T(0.5) * rL * qL) * Km1;
    const T cL = K * pL * irL;
    const T hL = T(0.5) * qL + cL * iKm1;

    const T rR = Rq[0], irR = T(1) / rR;
    const T uR = Rq[1] * irR, vR = Rq[2] * irR, wR = Rq[3] * irR;
    const T UnR = mx * uR + my * vR + mz * wR;
    const T qR = uR * uR + vR * vR + wR * wR;
    const T pR = (Rq[4] - T(0.5) * rR * qR) * Km1;
    const T cR = K * pR * irR;
    const T hR = T(0.5) * qR + cR * iKm1;

    const T sL = sqrt(rL), sR = sqrt(rR);
    const T iss = T(1) / (sL + sR);
    const T rho = sL * sR;
    const T u = (sL * uL + sR * uR) * iss;
    const T v = (sL * vL + sR * vR) * iss;
    const T w = (sL * wL + sR * wR) * iss;
    const T Un = (sL * UnL + sR * UnR) * iss;
    const T q = u * u + v * v + w * w;
    const T H = (sL * hL + sR * hR) * iss;
    const T C = (H - T(0.5) * q) * Km1;

    const T dU1 = rR - rL;
    const T dU2 = rR * uR - rL * uL;
    const T dU3 = rR * vR - rL * vL;
    const T dU4 = rR * wR - rL * wL;
    const T dU5 = (pR * iKm1 + T(0.5) * rR * qR) - (pL * iKm1 + T(0.5) * rL * qL);

    const T th = q / C;
    const T beta = (th > e) ? th : e;
    const T Unn = Un * insqr;
    const T ubar = T(0.5) * (T(1) + beta) * Unn;
    const T S = T(0.5) * sqrt(T(4) * beta * C + Unn * Unn * (T(1) - beta) * (T(1) - beta));
    const T iS = T(1) / S;
    const T U_p = T(0.5) * (T(1) + th) * Unn;
    const T C_p = T(0.5) * sqrt(T(4) * C * th + Unn * Unn * (T(1) - th) * (T(1) - th));

    const T t1 = (pR - pL) / (rho * beta * C);
    const T dUn = (UnR - UnL) * insqr;
    const T t2 = U_p * iS * dUn;
    const T t3 = T(0.5) * (T(1) - beta) * Unn * ubar * iS;
    const T t4 = T(0.5) * (T(1) - th) * Unn * U_p * iS;
    const T aU = fabs(Unn);
    const T dUc = (S - t3 - beta * aU) * t1 + t2;
    const T dPc = U_p * iS * (pR - pL) + (C_p - aU + t4) * rho * dUn;

    const T Fav1 = aU * dU1 + dUc * rho;
    const T Fav2 = aU * dU2 + dUc * rho * u + dPc * nx_;
    const T Fav3 = aU * dU3 + dUc * rho * v + dPc * ny_;
    const T Fav4 = aU * dU4 + dUc * rho * w + dPc * nz_;
    const T Fav5 = aU * dU5 + dUc * rho * H + dPc * Unn;

    const T hJ = T(0.5) / Jf, es = ep * sq;
    F[0] = ((rL * UnL + rR * UnR) - es * Fav1) * hJ;
    F[1] = ((rL * uL * UnL + rR * uR * UnR + mx * (pL + pR)) - es * Fav2) * hJ;
    F[2] = ((rL * vL * UnL + rR * vR * UnR + my * (pL + pR)) - es * Fav3) * hJ;
    F[3] = ((rL * wL * UnL + rR * wR * UnR + mz * (pL + pR)) - es * Fav4) * hJ;
    F[4] = ((UnL * (T(3.5) * pL + T(0.5) * rL * qL) + UnR * (T(3.5) * pR + T(0.5) * rR * qR)) - es * Fav5) * hJ;
    pL_out = pL;
    pR_out = pR;
    Fav3_out = Fav3;
}

// ---------------------------------------------------------------- x faces
// Face i+1/2 between cells (i, j+1, k+1) and (i+1, j+1, k+1), stored at
// [i][j][k] for i=2..gend, j=1..NY-1, k=1..NZ-1.  A thread owns a (j, k)
// column and marches i through a segment of faces: the six cells of the
// current face (planes i-2 .. i+3: conserved variables and 1/J) stay in
// registers and move up one plane per face, so each cell is loaded once per
// segment and each 1/J formed once; the next plane's cell and the next
// face's metrics are loaded a face ahead.  The same arithmetic as one
// thread per face (the window holds the values its loads would give).
template <typename T, bool SCALE>
__global__ void __launch_bounds__(256) k_flux_x(Geo g, double K, double e, C5 U, const double *__restrict__ J,
                                                const double *__restrict__ Jf, M9 mf, F5 F, double epS,
                                                const double *__restrict__ EpF, int seg)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int iS = 2 + blockIdx.z * seg, iE = min(iS + seg - 1, g.gend);
    if (k > g.NZ - 1 || j > g.NY - 1) return;
    const long sx = g.Ym * g.Zm;
    const long cc = ix(g, 0, j + 1, k + 1), fc = ix(g, 0, j, k);
    T wu[6][5], wq[6];        // planes i-2 .. i+3 of the current face i
    T nu[5], nJ;              // plane i+4, in flight
    T mx, my, mz, jf, ep;     // the current face's metrics
    T nmx, nmy, nmz, njf, nep;   // the next face's, in flight
    auto load_cell = [&](int pl, T u[5], T &Jv) {
#pragma unroll
        for (int v = 0; v < 5; v++) u[v] = ld<T>(U.v[v], cc + pl * sx);
        Jv = ld<T>(J, cc + pl * sx);
    };
    auto load_face = [&](int i, T &a, T &b, T &c, T &d, T &f) {
        const long fi = fc + i * sx;
        a = ld<T>(mf.m[0], fi);
        b = ld<T>(mf.m[1], fi);
        c = ld<T>(mf.m[2], fi);
        d = ld<T>(Jf, fi);
        f = SCALE ? (EpF ? ld<T>(EpF, fi) : T(epS)) : T(1);
    };
#pragma unroll
    for (int m = 0; m < 6; m++) {
        T Jv;
        load_cell(iS - 2 + m, wu[m], Jv);
        wq[m] = Jv;
    }
    load_face(iS, mx, my, mz, jf, ep);
#pragma unroll
    for (int m = 0; m < 6; m++) wq[m] = T(1) / wq[m];
    for (int i = iS; i <= iE; i++) {
        // a face ahead (the last face's loads repeat its own: in range, unused)
        load_cell(min(i + 4, iE + 3), nu, nJ);
        load_face(min(i + 1, iE), nmx, nmy, nmz, njf, nep);
        // left state of face i+1/2; later reference writes win (last-rank
        // patch after first-rank patch after the 5-point loop)
        T cl[6], cr[6];
        if (g.last && i == g.gend) sh3L<T>(cl);
        else if (g.first && i == 3) sh3L<T>(cl);
        else if (g.first && i == 2) sh23<T>(cl);
        else sh5L<T>(cl);
        // right state of face i+1/2 (from cell i+1)
        if (g.last && i == g.gend) sh23<T>(cr);
        else if (g.last && i == g.gend - 1) sh3R<T>(cr);
        else if (g.first && i == 2) sh3R<T>(cr);
        else sh5R<T>(cr);
        Sten<T> S;
#pragma unroll
        for (int m = 0; m < 6; m++) {
            S.q[m] = wq[m];
#pragma unroll
            for (int v = 0; v < 5; v++) S.u[v][m] = wu[m][v];
        }
        if ((g.first && i <= 3) || (g.last && i >= g.gend - 1)) {
            // a position neither stencil uses takes cell i's values (the
            // loads of one thread per face read cell i there)
#pragma unroll
            for (int m = 0; m < 6; m++)
                if (cl[m] == T(0) && cr[m] == T(0)) {
                    S.q[m] = wq[2];
#pragma unroll
                    for (int v = 0; v < 5; v++) S.u[v][m] = wu[2][v];
                }
        }
        const T wl = T(1) / dot6(cl, S.q), wr = T(1) / dot6(cr, S.q);
        T Lq[5], Rq[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            Lq[v] = wl * dot6(cl, S.u[v]);
            Rq[v] = wr * dot6(cr, S.u[v]);
        }
        const T insqr = T(1.0) / sqrt(mx * mx + my * my + mz * mz);
        T out[5], pl, pr, fav3;
        roe_face<T>(Lq, Rq, mx, my, mz, jf, T(e), T(K), mx * insqr, my * insqr, mz * insqr, ep, out, pl, pr, fav3);
        const long f = fc + i * sx;
#pragma unroll
        for (int v = 0; v < 5; v++) F.v[v][f] = (double)out[v];
        // move the window up one plane
#pragma unroll
        for (int m = 0; m < 5; m++) {
            wq[m] = wq[m + 1];
#pragma unroll
            for (int v = 0; v < 5; v++) wu[m][v] = wu[m + 1][v];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) wu[5][v] = nu[v];
        wq[5] = T(1) / nJ;
        mx = nmx;
        my = nmy;
        mz = nmz;
        jf = njf;
        ep = nep;
    }
}

// ---------------------------------------------------------------- y faces
// Face j+1/2 between cells (i+1, j, k+1) and (i+1, j+1, k+1), stored at
// [i][j][k] for i=2..gend-1, j=1..NY, k=1..NZ-1.  Near-wall order drop of
// Flux_Y.cpp:135-249 and the wall-face override of Flux_Y.cpp:400-408.  A
// thread owns an (i, k) column and marches j through a segment of faces
// with the six cells of the current face in registers, as the x faces do.
// Rows outside the array's j range are read from the nearest row (only
// where the stencil does not use them).
template <typename T, bool SCALE>
__global__ void __launch_bounds__(256) k_flux_y(Geo g, double K, double Ep, double e, int xi_mode, C5 U,
                                                const double *__restrict__ J, const double *__restrict__ Jf,
                                                M9 mf, F5 F, double epS, const double *__restrict__ EpF, int seg)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int i = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int jS = 1 + blockIdx.z * seg, jE = min(jS + seg - 1, g.NY);
    if (k > g.NZ - 1 || i > g.gend - 1) return;
    const int NY = g.NY, rmax = (int)g.Ym - 1;
    const long cc = ix(g, i + 1, 0, k + 1), fc = ix(g, i, 0, k);
    const long sy = g.Zm;
    T wu[6][5], wq[6];        // rows j-2 .. j+3 of the current face j
    T nu[5], nJ;              // row j+4, in flight
    T mx, my, mz, jf, ep, nmx, nmy, nmz, njf, nep;
    auto load_cell = [&](int r, T u[5], T &Jv) {
        const long c = cc + (long)min(max(r, 0), rmax) * sy;
#pragma unroll
        for (int v = 0; v < 5; v++) u[v] = ld<T>(U.v[v], c);
        Jv = ld<T>(J, c);
    };
    auto load_face = [&](int j, T &a, T &b, T &c, T &d, T &f) {
        const long fj = fc + (long)j * sy;
        a = ld<T>(mf.m[3], fj);
        b = ld<T>(mf.m[4], fj);
        c = ld<T>(mf.m[5], fj);
        d = ld<T>(Jf, fj);
        f = SCALE ? (EpF ? ld<T>(EpF, fj) : T(epS)) : T(1);
    };
#pragma unroll
    for (int m = 0; m < 6; m++) load_cell(jS - 2 + m, wu[m], wq[m]);
    load_face(jS, mx, my, mz, jf, ep);
    // 1/J where the row exists (a row outside it is never used)
#pragma unroll
    for (int m = 0; m < 6; m++) wq[m] = T(1) / wq[m];
    for (int j = jS; j <= jE; j++) {
        load_cell(j + 4, nu, nJ);
        load_face(min(j + 1, jE), nmx, nmy, nmz, njf, nep);
        T cl[6], cr[6];
        if (j == NY) sh12<T>(cl);
        else if (j == NY - 1) sh3L<T>(cl);
        else if (j == 3) sh3L<T>(cl);
        else if (j == 2) sh23<T>(cl);
        else if (j == 1) sh23<T>(cl);
        else sh5L<T>(cl);
        if (j == NY) sh23<T>(cr);
        else if (j == NY - 1) sh23<T>(cr);
        else if (j == NY - 2) sh3R<T>(cr);
        else if (j == 2) sh3R<T>(cr);
        else if (j == 1) sh34<T>(cr);
        else sh5R<T>(cr);
        Sten<T> S;
#pragma unroll
        for (int m = 0; m < 6; m++) {
            S.q[m] = wq[m];
#pragma unroll
            for (int v = 0; v < 5; v++) S.u[v][m] = wu[m][v];
        }
        if (j <= 3 || j >= NY - 2) {
            // a position neither stencil uses takes cell j's values (the
            // loads of one thread per face read cell j there)
#pragma unroll
            for (int m = 0; m < 6; m++)
                if (cl[m] == T(0) && cr[m] == T(0)) {
                    S.q[m] = wq[2];
#pragma unroll
                    for (int v = 0; v < 5; v++) S.u[v][m] = wu[2][v];
                }
        }
        const T wl = T(1) / dot6(cl, S.q), wr = T(1) / dot6(cr, S.q);
        T Lq[5], Rq[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            Lq[v] = wl * dot6(cl, S.u[v]);
            Rq[v] = wr * dot6(cr, S.u[v]);
        }
        const T insqr = T(1.0) / sqrt(mx * mx + my * my + mz * mz);
        T nx_ = 0, ny_ = 0, nz_ = 0;
        if (xi_mode) { nx_ = mx * insqr; ny_ = my * insqr; nz_ = mz * insqr; }
        T out[5], pl, pr, fav3;
        roe_face<T>(Lq, Rq, mx, my, mz, jf, T(e), T(K), nx_, ny_, nz_, ep, out, pl, pr, fav3);
        if (j == 1 || j == NY) {
            out[0] = 0;
            out[1] = 0;
            out[2] = T(0.5) * (my * (pl + pr) - T(Ep) * fav3) / jf;
            out[3] = 0;
            out[4] = 0;
        }
        const long f = fc + (long)j * sy;
#pragma unroll
        for (int v = 0; v < 5; v++) F.v[v][f] = (double)out[v];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            wq[m] = wq[m + 1];
#pragma unroll
            for (int v = 0; v < 5; v++) wu[m][v] = wu[m + 1][v];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) wu[5][v] = nu[v];
        wq[5] = T(1) / nJ;
        mx = nmx;
        my = nmy;
        mz = nmz;
        jf = njf;
        ep = nep;
    }
}

// ---------------------------------------------------------------- z faces
// Face k+1/2 between cells (i+1, j+1, k) and (i+1, j+1, k+1), stored at
// [i][j][k] for i=2..gend-1, j=1..NY-1, k=1..NZ (Flux_Z.cpp:105-216).
template <typename T, bool SCALE>
__global__ void __launch_bounds__(256) k_flux_z(Geo g, double K, double e, int xi_mode, C5 U,
                                                const double *__restrict__ J, const double *__restrict__ Jf,
                                                M9 mf, F5 F, double epS, const double *__restrict__ EpF)
{
    // (j, k) = (1..NY-1, 1..NZ) of a plane flattened, k fastest: NZ faces
    // per row would leave a third of the lanes idle in 64-wide rows
    const int nk = g.NZ;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * (g.NY - 1)) return;
    const int k = 1 + t % nk;
    const int j = 1 + t / nk;
    const int i = 2 + blockIdx.z;
    const long c0 = ix(g, i + 1, j + 1, k);   // cell k
    const int NZ = g.NZ;
    T cl[6], cr[6];
    if (k == NZ) sh3L<T>(cl);
    else if (k == 2) sh3L<T>(cl);
    else if (k == 1) sh23<T>(cl);
    else sh5L<T>(cl);
    if (k == NZ) sh23<T>(cr);
    else if (k == NZ - 1) sh3R<T>(cr);
    else if (k == 1) sh3R<T>(cr);
    else sh5R<T>(cr);
    Sten<T> S;
    load_sten<T>(U, J, c0, 1, cl, cr, S);
    const long f = ix(g, i, j, k);
    const T ztx = ld<T>(mf.m[6], f), zty = ld<T>(mf.m[7], f), ztz = ld<T>(mf.m[8], f);
    const T jf = ld<T>(Jf, f);
    const T ep = SCALE ? (EpF ? ld<T>(EpF, f) : T(epS)) : T(1);
    // every load of the face is in flight before the first waits
    __builtin_amdgcn_sched_barrier(0);
    T Lq[5], Rq[5];
    face_states<T>(S, cl, cr, Lq, Rq);
    const T insqr = T(1.0) / (sqrt(ztx * ztx + zty * zty + ztz * ztz));
    T nx_ = 0, ny_ = 0, nz_ = 0;
    if (xi_mode) { nx_ = ztx * insqr; ny_ = zty * insqr; nz_ = ztz * insqr; }
    T out[5], pl, pr, fav3;
    roe_face<T>(Lq, Rq, ztx, zty, ztz, jf, T(e), T(K), nx_, ny_, nz_, ep, out, pl, pr, fav3);
#pragma unroll
    for (int v = 0; v < 5; v++) F.v[v][f] = (double)out[v];
}

// the unscaled instantiation is the live reference flux, compiled without
// the factor
#define SCALED(K, T) ((EpF || epS != 1.0) ? K<T, true> : K<T, false>)

static dim3 grid_for(int nk, int nj, int ni, dim3 b)
{
    return dim3((nk + b.x - 1) / b.x, (nj + b.y - 1) / b.y, ni);
}

void launch_flux_x(const Geo &g, const Phys &p, double e, C5 U, const double *J, const double *Jf, M9 mf,
                   F5 F, double epS, const double *EpF, hipStream_t s)
{
    // a block spans the whole k row where it fits (255 faces at NZ = 256)
    const int bx = g.NZ - 1 > 128 ? 256 : g.NZ - 1 > 64 ? 128 : 64;
    dim3 b(bx, 256 / bx);
    // faces i = 2..gend in segments: about three waves per SIMD in one round
    const int nf = g.gend - 1;
    const long cols = (long)((g.NZ - 1 + bx - 1) / bx) * ((g.NY - 1 + b.y - 1) / b.y) * 4;   // waves per segment row
    const int nseg = (int)std::max(1L, std::min((long)nf, 3072 / std::max(1L, cols)));
    const int seg = (nf + nseg - 1) / nseg;
    const dim3 gr = grid_for(g.NZ - 1, g.NY - 1, (nf + seg - 1) / seg, b);
    if (p.fp32)
        hipLaunchKernelGGL(SCALED(k_flux_x, float), gr, b, 0, s, g, p.K, e, U, J, Jf, mf, F, epS, EpF, seg);
    else
        hipLaunchKernelGGL(SCALED(k_flux_x, double), gr, b, 0, s, g, p.K, e, U, J, Jf, mf, F, epS, EpF, seg);
}

void launch_flux_y(const Geo &g, const Phys &p, double Ep, double e, int xi_mode, C5 U, const double *J,
                   const double *Jf, M9 mf, F5 F, double epS, const double *EpF, hipStream_t s)
{
    const int bx = g.NZ - 1 > 128 ? 256 : g.NZ - 1 > 64 ? 128 : 64;
    dim3 b(bx, 256 / bx);
    // faces j = 1..NY in segments: about three waves per SIMD in one round
    const int nf = g.NY;
    const long cols = (long)((g.NZ - 1 + bx - 1) / bx) * ((g.gend - 2 + b.y - 1) / b.y) * 4;
    const int nseg = (int)std::max(1L, std::min((long)nf, 3072 / std::max(1L, cols)));
    const int seg = (nf + nseg - 1) / nseg;
    const dim3 gr = grid_for(g.NZ - 1, g.gend - 2, (nf + seg - 1) / seg, b);
    if (p.fp32)
        hipLaunchKernelGGL(SCALED(k_flux_y, float), gr, b, 0, s, g, p.K, Ep, e, xi_mode, U, J, Jf, mf, F, epS, EpF,
                           seg);
    else
        hipLaunchKernelGGL(SCALED(k_flux_y, double), gr, b, 0, s, g, p.K, Ep, e, xi_mode, U, J, Jf, mf, F, epS, EpF,
                           seg);
}

void launch_flux_z(const Geo &g, const Phys &p, double e, int xi_mode, C5 U, const double *J, const double *Jf,
                   M9 mf, F5 F, double epS, const double *EpF, hipStream_t s)
{
    const dim3 b(256);
    const dim3 gr((g.NZ * (g.NY - 1) + 255) / 256, 1, g.gend - 2);
    if (p.fp32)
        hipLaunchKernelGGL(SCALED(k_flux_z, float), gr, b, 0, s, g, p.K, e, xi_mode, U,
                           J, Jf, mf, F, epS, EpF);
    else
        hipLaunchKernelGGL(SCALED(k_flux_z, double), gr, b, 0, s, g, p.K, e, xi_mode, U,
                           J, Jf, mf, F, epS, EpF);
}

}  // namespace cgli
