// filter.hip -- the multi-scale filter sensor (Filter.cpp:12-1720) on gfx950.
//
// Two scales (h: 1/8 (1,6,1); 2h: 1/4 (1,2,1)), six levels each of the
// separable filter G = Gz Gy Gx applied to U1..U4 (U5 is filtered by the
// reference but never used).  One launch per pass:
//   k_filt_x    Gx = x pass of the level input      (Filter.cpp:162, 841)
//   k_filt_y    Gy = J/Y_point-weighted y pass of Gx (Filter.cpp:199, 875)
//   k_filt_z    level output = z pass of Gy         (Filter.cpp:231, 902)
//   k_filt_exch GG halo, periodic single slab       (Filter.cpp:433-466)
//   k_filt_energy 6G-15G2+20G3-15G4+6G5-G6, then E = |u - u_f|^2
//                 (Filter.cpp:769-800, 1444-1475)
//   k_filt_eghost ghost cells of E copy the interior (Filter.cpp:1480-1548)
//   k_filt_sensor EpX/EpY/EpZ = E_high or E_low      (Filter.cpp:1560-1645)
// Expressions keep the reference's operation order, so results are
// bit-identical with the reference built by g++ on x86 (no FMA
// contraction: the library is compiled with -ffp-contract=off).
// Bytes per pass: 4 fields in, 4 out (the y pass adds J and Y_point).
// Every kernel is templated on the arithmetic type Real: FP64, or FP32 for
// the mixed-precision path (work arrays, passes, energies and the sensor
// ratio in float; U_, J and Y_point read as FP64 and rounded once; the
// E_high/E_low outputs are FP64 fields either way).
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

// every work array starts as U_ on i=0..gend+3, j=1..NYY, k=1..NZZ
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
        for (int a = 0; a < 8; a++) work[(a * 4 + v) * N + c] = u;
    }
}

// In: the level's input type (U_ itself, FP64, at level 0)
template <bool TWOH, typename Real, typename In>
__global__ void k_filt_x(Geo g, int lo, R4T<In> in, L4T<Real> out)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = lo + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k), sx = g.Ym * g.Zm;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const In *u = in.v[v];
        const Real um = (Real)u[c - sx], u0 = (Real)u[c], up = (Real)u[c + sx];
        if (!TWOH)
            out.v[v][c] = Real(0.125) * (um + 6 * u0 + up);
        else
            out.v[v][c] = Real(0.25) * (um + 2 * u0 + up);
    }
}

template <bool TWOH, typename Real>
__global__ void k_filt_y(Geo g, int lo, R4T<Real> gx, L4T<Real> gy, const double *__restrict__ J,
                         const double *__restrict__ Y)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = lo + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k), jm = c - g.Zm, jp = c + g.Zm;
    const Real a = (Real)(Y[jp] - Y[jm]);
    const Real b = (Real)(Y[jp] - Y[c]);
    const Real cc = (Real)(Y[c] - Y[jm]);
    const Real Jm = (Real)J[jm], Jc = (Real)J[c], Jp = (Real)J[jp];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const Real *x = gx.v[v];
        if (!TWOH)
            gy.v[v][c] = Real(0.125) * (2 * b / a * x[jm] * Jm + 6 * x[c] * Jc + 2 * cc / a * x[jp] * Jp) / Jc;
        else
            gy.v[v][c] = Real(0.5) * (b / a * x[jm] * Jm + x[c] * Jc + cc / a * x[jp] * Jp) / Jc;
    }
}

template <bool TWOH, typename Real>
__global__ void k_filt_z(Geo g, int lo, R4T<Real> gy, L4T<Real> out)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = lo + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const Real *y = gy.v[v];
        if (!TWOH)
            out.v[v][c] = Real(0.125) * (y[c - 1] + 6 * y[c] + y[c + 1]);
        else
            out.v[v][c] = Real(0.25) * (y[c - 1] + 2 * y[c] + y[c + 1]);
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

long filter_scratch_doubles(const Geo &g) { return 34 * g.N; }

// work (34 N doubles; the FP32 path uses the first half as 34 N floats):
// levels 0..5 x 4 variables, Gx (4), Gy (4), E_h, E_2h
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
    const L4T<Real> Gx = lvl(6), Gy = lvl(7);
    Real *Eh = work + 32 * N, *E2h = work + 33 * N;
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_filt_init<Real>, dim3((g.NZZ + 63) / 64, (g.NYY + 3) / 4, g.gend + 4), b, 0, s, g, U, work,
                       N);
    for (int sc = 0; sc < 2; sc++) {
        for (int l = 0; l < 6; l++) {
            int lo, hi;
            level_range(g, l, lo, hi);
            dim3 gr((g.NZ - 1 + 63) / 64, (g.NY - 1 + 3) / 4, hi - lo + 1);
            if (sc == 0) {
                if (l == 0)
                    hipLaunchKernelGGL((k_filt_x<false, Real, double>), gr, b, 0, s, g, lo, U, Gx);
                else
                    hipLaunchKernelGGL((k_filt_x<false, Real, Real>), gr, b, 0, s, g, lo, cst(lvl(l - 1)), Gx);
                hipLaunchKernelGGL((k_filt_y<false, Real>), gr, b, 0, s, g, lo, cst(Gx), Gy, J, Y);
                hipLaunchKernelGGL((k_filt_z<false, Real>), gr, b, 0, s, g, lo, cst(Gy), lvl(l));
            } else {
                if (l == 0)
                    hipLaunchKernelGGL((k_filt_x<true, Real, double>), gr, b, 0, s, g, lo, U, Gx);
                else
                    hipLaunchKernelGGL((k_filt_x<true, Real, Real>), gr, b, 0, s, g, lo, cst(lvl(l - 1)), Gx);
                hipLaunchKernelGGL((k_filt_y<true, Real>), gr, b, 0, s, g, lo, cst(Gx), Gy, J, Y);
                hipLaunchKernelGGL((k_filt_z<true, Real>), gr, b, 0, s, g, lo, cst(Gy), lvl(l));
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
