// filter.hip -- the multi-scale filter sensor (Filter.cpp:12-1720) on gfx950.
//
// Two scales (h: 1/8 (1,6,1); 2h: 1/4 (1,2,1)), six levels each of the
// separable filter G = Gz Gy Gx applied to U1..U4 (U5 is filtered by the
// reference but never used).  Launches:
//   k_filt_level  one level: Gx = x pass of the level input, Gy =
//                 J/Y_point-weighted y pass of Gx, output = z pass of Gy,
//                 fused (Filter.cpp:162/199/231, 841/875/902)
//   k_filt_exch GG halo, periodic single slab       (Filter.cpp:433-466)
//   k_filt_energy 6G-15G2+20G3-15G4+6G5-G6, then E = |u - u_f|^2
//                 (Filter.cpp:769-800, 1444-1475)
//   k_filt_eghost ghost cells of E copy the interior (Filter.cpp:1480-1548)
//   k_filt_sensor EpX/EpY/EpZ = E_high or E_low      (Filter.cpp:1560-1645)
// Expressions keep the reference's operation order, so results are
// bit-identical with the reference built by g++ on x86 (no FMA
// contraction: the library is compiled with -ffp-contract=off).
// Bytes per level: 4 fields in (plus a 1-cell halo in j and k), J and
// Y_point, 4 fields out.  Segments are short (about 4096 blocks per level
// launch): more blocks in flight hide the plane-to-plane latency better than
// the halo planes each segment re-reads cost.
// Every kernel is templated on the arithmetic type Real: FP64, or FP32 for
// the mixed-precision path (work arrays, passes, energies and the sensor
// ratio in float; U_, J and Y_point read as FP64 and rounded once; the
// E_high/E_low outputs are FP64 fields either way).
#include <algorithm>
#include <type_traits>

#include "cgli_internal.hpp"

namespace cgli {

template <typename T> struct L4T { T *v[4]; };
template <typename T> struct R4T { const T *v[4]; };
typedef L4T<double> L4;
typedef R4T<double> R4;

// x range of level l (0..5): Filter.cpp:149-153 (levels 0, 3), 253-257
// (levels 1, 4), 348-349 (levels 2, 5)
static void level_range(const Geo &g, int lvl, int &lo, int &hi)
{
    const int w = (lvl == 0 || lvl == 3) ? 2 : ((lvl == 1 || lvl == 4) ? 1 : 0);
    lo = g.first ? 3 : 3 - w;
    hi = g.last ? g.gend : g.gend + w;
}

// every level array starts as U_ on i=0..gend+3, j=1..NYY, k=1..NZZ
// (Filter.cpp:117-134).  Only the box's ghost shell (the three x planes at
// each end, j=1 and j=NYY, k=1 and k=NZZ) is read before the passes
// overwrite it, so only the shell is copied: the interior would be
// overwritten unread.  (The work arrays start as NaN when allocated, so a
// read of an unwritten interior cell would spoil the sensor.)
template <typename Real>
__global__ void k_filt_init(Geo g, R4 U, Real *__restrict__ work, long N)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k > g.NZZ || j > g.NYY) return;
    const bool shell = i <= 2 || i >= g.gend + 1 || j == 1 || j == g.NYY || k == 1 || k == g.NZZ;
    if (!shell) return;
    const long c = ix(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const Real u = (Real)U.v[v][c];
#pragma unroll
        for (int a = 0; a < 6; a++) work[(a * 4 + v) * N + c] = u;
    }
}

// One level of one scale (Filter.cpp:162-257 and 841-902 for the two
// scales), its three passes fused:
//   Gx = x pass of the level input         1/8 (1,6,1) or 1/4 (1,2,1)
//   Gy = J/Y_point-weighted y pass of Gx
//   level output = z pass of Gy
// on j = 2..NY, k = 2..NZ and the level's planes lo..hi.  A block owns an
// 4 x 64 tile of (j, k) and marches along i through a segment of planes.
// Each thread keeps the level input of its halo positions (the tile plus a
// 1-cell halo) for planes i-1, i, i+1 in registers, with plane i+2 in
// flight, so each plane is read from HBM once per block and the x pass
// needs no LDS; Gx and Gy of plane i go through LDS for the y and z passes,
// and the level output is the only store.  Where the reference reads Gx on
// the ghost rows j = 1, NYY or Gy on the ghost columns k = 1, NZZ, which no
// pass writes and which hold U_ from its initial copy (Filter.cpp:117-134),
// this kernel reads U_: those positions load U_ in place of the level
// input.  Expressions and operation order are the reference's; the y
// pass's quotients that share a divisor use its reciprocal and div_r, which
// gives the division's bits.
constexpr int FT_J = 4, FT_K = 64, FT_THREADS = FT_J * FT_K;
constexpr int FH_K = FT_K + 2, FH_N = (FT_J + 2) * FH_K;

template <bool TWOH, typename Real, typename In>
__global__ void __launch_bounds__(FT_THREADS) k_filt_level(Geo g, int lo, int hi, int seg, R4T<In> prev, R4 U,
                                                          const double *__restrict__ J, const double *__restrict__ Y,
                                                          L4T<Real> out)
{
    constexpr int NR = (FH_N + FT_THREADS - 1) / FT_THREADS;   // halo'd positions per thread
    __shared__ Real sX[4][FH_N];      // Gx of plane i (U_ on the ghost rows and columns)
    __shared__ double sJ[FH_N], sY[FH_N];
    __shared__ Real sG[4][FT_J][FH_K];   // Gy of plane i, columns k0-1 .. k0+FT_K
    const int j0 = 2 + blockIdx.y * FT_J, k0 = 2 + blockIdx.x * FT_K;
    const int iS = lo + blockIdx.z * seg, iE = min(iS + seg - 1, hi);
    if (iS > hi) return;   // the whole block
    // this thread's halo'd tile positions t = threadIdx.x + r FT_THREADS:
    // (j0 - 1 + t / FH_K, k0 - 1 + t % FH_K), where inside the arrays' box
    long cpos[NR];
    bool inb[NR], ghost[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int t = threadIdx.x + r * FT_THREADS;
        const int j = j0 - 1 + t / FH_K, k = k0 - 1 + t % FH_K;
        inb[r] = t < FH_N && j <= g.NYY && k <= g.NZZ;
        ghost[r] = j == 1 || j == g.NYY || k == 1 || k == g.NZZ;
        cpos[r] = inb[r] ? ix(g, 0, j, k) : 0;
    }
    const long sx = g.Ym * g.Zm;
    // The x pass runs down each position's own column: the level input of
    // planes i-1, i, i+1 in registers (xm, x0, xp), plane i+2 in flight
    // (pL), so the x pass needs no LDS.  A ghost row or column holds U_
    // instead (the passes read U_ there).  J and Y_point of a plane are
    // loaded into registers one iteration before they go to LDS.  Every
    // load runs for every slot (a position outside the box reads plane i's
    // first cell and is not stored): no branch around them.
    Real xm[NR][4], x0[NR][4], xp[NR][4], pL[NR][4];
    double pJ[NR], pY[NR];
    auto fetch = [&](Real (*dst)[4], int i) {
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const long c = cpos[r] + i * sx;
                if constexpr (std::is_same<In, double>::value) {
                    const double *src = ghost[r] ? U.v[v] : prev.v[v];
                    dst[r][v] = (Real)src[c];
                } else {
                    const Real a = (Real)prev.v[v][c], b = (Real)U.v[v][c];
                    dst[r][v] = ghost[r] ? b : a;
                }
            }
    };
    auto fetch_JY = [&](int i) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            pJ[r] = J[cpos[r] + i * sx];
            pY[r] = Y[cpos[r] + i * sx];
        }
    };
    auto put_JY = [&]() {
#pragma unroll
        for (int r = 0; r < NR; r++)
            if (inb[r]) {
                sJ[threadIdx.x + r * FT_THREADS] = pJ[r];
                sY[threadIdx.x + r * FT_THREADS] = pY[r];
            }
    };
    fetch(xm, iS - 1);
    fetch(x0, iS);
    fetch(xp, iS + 1);
    fetch_JY(iS);
    put_JY();
    fetch(pL, min(iS + 2, iE + 1));
    fetch_JY(min(iS + 1, iE));
    for (int i = iS; i <= iE; i++) {
        __syncthreads();
        // Gx on rows j0-1 .. j0+FT_J, columns 1..NZZ of the halo'd tile
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int t = threadIdx.x + r * FT_THREADS;
            if (!inb[r]) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (ghost[r])
                    sX[v][t] = x0[r][v];
                else if (!TWOH)
                    sX[v][t] = Real(0.125) * (xm[r][v] + 6 * x0[r][v] + xp[r][v]);
                else
                    sX[v][t] = Real(0.25) * (xm[r][v] + 2 * x0[r][v] + xp[r][v]);
            }
        }
        __syncthreads();
        // Gy on rows j0 .. j0+FT_J-1 (up to NY), columns k0-1 .. k0+FT_K
        for (int t = threadIdx.x; t < FT_J * FH_K; t += FT_THREADS) {
            const int tj = t / FH_K, hk = t % FH_K;
            const int j = j0 + tj, k = k0 - 1 + hk;
            if (j > g.NY || k > g.NZZ) continue;
            const int pc = (tj + 1) * FH_K + hk, pm = pc - FH_K, pp = pc + FH_K;
            if (k == 1 || k == g.NZZ) {
#pragma unroll
                for (int v = 0; v < 4; v++) sG[v][tj][hk] = sX[v][pc];
                continue;
            }
            const Real a = (Real)(sY[pp] - sY[pm]);
            const Real b = (Real)(sY[pp] - sY[pc]);
            const Real cc = (Real)(sY[pc] - sY[pm]);
            const Real Jm = (Real)sJ[pm], Jc = (Real)sJ[pc], Jp = (Real)sJ[pp];
            // the quotients by a and by Jc through their reciprocals (div_r:
            // the division's bits)
            const Real ia = Real(1) / a, iJc = Real(1) / Jc;
            const Real wm = div_r(TWOH ? b : 2 * b, a, ia), wp = div_r(TWOH ? cc : 2 * cc, a, ia);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const Real *x = sX[v];
                if (!TWOH)
                    sG[v][tj][hk] = div_r(Real(0.125) * (wm * x[pm] * Jm + 6 * x[pc] * Jc + wp * x[pp] * Jp), Jc, iJc);
                else
                    sG[v][tj][hk] = div_r(Real(0.5) * (wm * x[pm] * Jm + x[pc] * Jc + wp * x[pp] * Jp), Jc, iJc);
            }
        }
        __syncthreads();
        // J and Y_point of plane i are read: they take those of plane i+1;
        // the column window moves up one plane
        put_JY();
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                xm[r][v] = x0[r][v];
                x0[r][v] = xp[r][v];
                xp[r][v] = pL[r][v];
            }
        fetch(pL, min(i + 3, iE + 1));
        fetch_JY(min(i + 2, iE));
        {   // the level output: z pass of Gy
            const int tj = threadIdx.x / FT_K, tk = threadIdx.x % FT_K;
            const int j = j0 + tj, k = k0 + tk;
            if (j <= g.NY && k <= g.NZ) {
                const long c = ix(g, i, j, k);
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const Real *y = sG[v][tj];
                    if (!TWOH)
                        out.v[v][c] = Real(0.125) * (y[tk] + 6 * y[tk + 1] + y[tk + 2]);
                    else
                        out.v[v][c] = Real(0.25) * (y[tk] + 2 * y[tk + 1] + y[tk + 2]);
                }
            }
        }
    }
}

// periodic image of the three planes at each x end
template <typename Real>
__global__ void k_filt_exch(Geo g, L4T<Real> G)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int q = blockIdx.z;   // 0..2
    if (k > g.NZZ || j > g.NYY) return;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        Real *a = G.v[v];
        a[ix(g, g.gend + 1 + q, j, k)] = a[ix(g, 3 + q, j, k)];
        a[ix(g, q, j, k)] = a[ix(g, g.gend - 2 + q, j, k)];
    }
}

// work: level l, variable v at (l*4 + v)*N; E over i=3..gend
template <typename Real>
__global__ void k_filt_energy(Geo g, R4 U, const Real *__restrict__ work, long N, Real *__restrict__ E)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
    Real f[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const Real *w = work + v * N + c;
        f[v] = 6 * w[0] - 15 * w[4 * N] + 20 * w[8 * N] - 15 * w[12 * N] + 6 * w[16 * N] - w[20 * N];
    }
    const Real U1 = (Real)U.v[0][c];
    const Real Ue = (Real)U.v[1][c] / U1 - f[1] / f[0];
    const Real Ve = (Real)U.v[2][c] / U1 - f[2] / f[0];
    const Real We = (Real)U.v[3][c] / U1 - f[3] / f[0];
    E[c] = Ue * Ue + Ve * Ve + We * We;
}

// ghost cells of E (both scales): x ends when first/last (periodic images
// for a periodic slab), then the j and k ghost layers of the interior i
template <typename Real>
__global__ void k_filt_eghost(Geo g, Real *__restrict__ Eh, Real *__restrict__ E2h)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = g.NY - 1, nz = g.NZ - 1, nxi = g.nx;
    // part 1: x ends over (j,k) in 2..NY x 2..NZ
    const int nx_part = ny * nz;
    // part 2: y rows over (i,k) in 3..gend x 2..NZ
    const int ny_part = nxi * nz;
    // part 3: z columns over (i,j) in 3..gend x 2..NY
    const int nz_part = nxi * ny;
    if (t < nx_part) {
        const int j = 2 + t / nz, k = 2 + t % nz;
        // an interior slab of a split domain: the reference reads planes
        // it never wrote there; they are zero here
        const long lo = ix(g, 2, j, k), hi = ix(g, g.gend + 1, j, k);
        if (g.first || g.xper) {
            const long src = ix(g, (g.xper && !g.first) ? g.gend : 3, j, k);
            Eh[lo] = Eh[src];
            E2h[lo] = E2h[src];
        } else {
            Eh[lo] = Real(0);
            E2h[lo] = Real(0);
        }
        if (g.last || g.xper) {
            const long src = ix(g, (g.xper && !g.last) ? 3 : g.gend, j, k);
            Eh[hi] = Eh[src];
            E2h[hi] = E2h[src];
        } else {
            Eh[hi] = Real(0);
            E2h[hi] = Real(0);
        }
    } else if (t < nx_part + ny_part) {
        const int u = t - nx_part;
        const int i = 3 + u / nz, k = 2 + u % nz;
        Eh[ix(g, i, 1, k)] = Eh[ix(g, i, 2, k)];
        Eh[ix(g, i, g.NYY, k)] = Eh[ix(g, i, g.NY, k)];
        E2h[ix(g, i, 1, k)] = E2h[ix(g, i, 2, k)];
        E2h[ix(g, i, g.NYY, k)] = E2h[ix(g, i, g.NY, k)];
    } else if (t < nx_part + ny_part + nz_part) {
        const int u = t - nx_part - ny_part;
        const int i = 3 + u / ny, j = 2 + u % ny;
        Eh[ix(g, i, j, 1)] = Eh[ix(g, i, j, 2)];
        Eh[ix(g, i, j, g.NZZ)] = Eh[ix(g, i, j, g.NZ)];
        E2h[ix(g, i, j, 1)] = E2h[ix(g, i, j, 2)];
        E2h[ix(g, i, j, g.NZZ)] = E2h[ix(g, i, j, g.NZ)];
    }
}

// one thread per (i,j,k) of the box i=2..gend, j=1..NY, k=1..NZ; each face
// family writes where its reference loop does
template <typename Real>
__global__ void k_filt_sensor(Geo g, const Real *__restrict__ Eh, const Real *__restrict__ E2h, Real crit,
                              double E_high, double E_low, double *__restrict__ EpX, double *__restrict__ EpY,
                              double *__restrict__ EpZ)
{
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 2 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k), sx = g.Ym * g.Zm, sy = g.Zm;
    if (j < g.NY && k < g.NZ) {   // Filter.cpp:1564-1583
        const long a = c + sy + 1, b = a + sx;
        const Real Emax = (Eh[a] + Eh[b]) / (E2h[a] + E2h[b]);
        EpX[c] = (Emax >= crit) ? E_high : E_low;
    }
    if (i <= g.gend - 1 && k < g.NZ) {   // Filter.cpp:1593-1614
        const long a = c + sx + 1, b = a + sy;
        const Real Emax = (Eh[a] + Eh[b]) / (E2h[a] + E2h[b]);
        EpY[c] = (Emax >= crit) ? E_high : E_low;
    }
    if (i <= g.gend - 1 && j < g.NY) {   // Filter.cpp:1624-1643
        const long a = c + sx + sy, b = a + 1;
        const Real Emax = (Eh[a] + Eh[b]) / (E2h[a] + E2h[b]);
        EpZ[c] = (Emax >= crit) ? E_high : E_low;
    }
}

long filter_scratch_doubles(const Geo &g) { return 26 * g.N; }

// work (26 N doubles; the FP32 path uses the first half as 26 N floats):
// levels 0..5 x 4 variables, E_h, E_2h
template <typename Real>
static void filter_run(const Geo &g, double crit, double E_high, double E_low, C5 U_, const double *J,
                       const double *Y, Real *work, double *EpX, double *EpY, double *EpZ, hipStream_t s)
{
    const long N = g.N;
    const R4 U{{U_.v[0], U_.v[1], U_.v[2], U_.v[3]}};
    auto lvl = [&](int l) {
        return L4T<Real>{{work + (l * 4 + 0) * N, work + (l * 4 + 1) * N, work + (l * 4 + 2) * N,
                          work + (l * 4 + 3) * N}};
    };
    auto cst = [](const L4T<Real> &a) { return R4T<Real>{{a.v[0], a.v[1], a.v[2], a.v[3]}}; };
    Real *Eh = work + 24 * N, *E2h = work + 25 * N;
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_filt_init<Real>, dim3((g.NZZ + 63) / 64, (g.NYY + 3) / 4, g.gend + 4), b, 0, s, g, U, work,
                       N);
    const int tk = (g.NZ - 1 + FT_K - 1) / FT_K, tj = (g.NY - 1 + FT_J - 1) / FT_J;
    for (int sc = 0; sc < 2; sc++) {
        for (int l = 0; l < 6; l++) {
            int lo, hi;
            level_range(g, l, lo, hi);
            // segments of at least 8 planes, about 4096 blocks
            const int np = hi - lo + 1;
            const int seg = std::min(np, std::max(8, (int)((long)np * tk * tj / 4096)));
            const dim3 gr(tk, tj, (np + seg - 1) / seg), bl(FT_THREADS);
            if (sc == 0) {
                if (l == 0)
                    hipLaunchKernelGGL((k_filt_level<false, Real, double>), gr, bl, 0, s, g, lo, hi, seg, U, U, J, Y,
                                       lvl(l));
                else
                    hipLaunchKernelGGL((k_filt_level<false, Real, Real>), gr, bl, 0, s, g, lo, hi, seg, cst(lvl(l - 1)), U,
                                       J, Y, lvl(l));
            } else {
                if (l == 0)
                    hipLaunchKernelGGL((k_filt_level<true, Real, double>), gr, bl, 0, s, g, lo, hi, seg, U, U, J, Y,
                                       lvl(l));
                else
                    hipLaunchKernelGGL((k_filt_level<true, Real, Real>), gr, bl, 0, s, g, lo, hi, seg, cst(lvl(l - 1)), U,
                                       J, Y, lvl(l));
            }
            if (l == 2 && g.xper)
                hipLaunchKernelGGL(k_filt_exch<Real>, dim3((g.NZZ + 63) / 64, (g.NYY + 3) / 4, 3), b, 0, s, g, lvl(2));
        }
        hipLaunchKernelGGL(k_filt_energy<Real>, dim3((g.NZ - 1 + 63) / 64, (g.NY - 1 + 3) / 4, g.nx), b, 0, s, g, U,
                           work, N, sc == 0 ? Eh : E2h);
    }
    const int nt = (g.NY - 1) * (g.NZ - 1) + g.nx * (g.NZ - 1) + g.nx * (g.NY - 1);
    hipLaunchKernelGGL(k_filt_eghost<Real>, dim3((nt + 255) / 256), dim3(256), 0, s, g, Eh, E2h);
    hipLaunchKernelGGL(k_filt_sensor<Real>, dim3((g.NZ + 63) / 64, (g.NY + 3) / 4, g.gend - 1), b, 0, s, g, Eh, E2h,
                       (Real)crit, E_high, E_low, EpX, EpY, EpZ);
}

void launch_filter(const Geo &g, double crit, double E_high, double E_low, C5 U_, const double *J,
                   const double *Y, double *work, int fp32, double *EpX, double *EpY, double *EpZ, hipStream_t s)
{
    if (fp32)
        filter_run<float>(g, crit, E_high, E_low, U_, J, Y, reinterpret_cast<float *>(work), EpX, EpY, EpZ, s);
    else
        filter_run<double>(g, crit, E_high, E_low, U_, J, Y, work, EpX, EpY, EpZ, s);
}

}  // namespace cgli
