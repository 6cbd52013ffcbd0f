// bc.hip -- ghost-cell fills on gfx950.
//   k_bc_y          Y_boundary_condition.cpp:98-182 (isothermal no-slip walls)
//   k_bc_z          Z_boundary_condition.cpp:48-66 (one periodic layer)
//   k_bc_x_periodic three periodic planes at each x end of a single slab
//   k_pack_q / k_unpack_q  one block of the restart file (Output.cpp:183-201,
//                   Initial_condition.cpp:196-212)
#include "cgli_internal.hpp"

namespace cgli {

__global__ void k_bc_y(Geo g, double K, double R, int gstart, F5 U, const double *__restrict__ J, double T_low,
                       double T_high, int heated_lo, int heated_hi)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int side = threadIdx.y;   // 0 bottom wall, 1 top wall
    const int i = 3 + blockIdx.z;
    if (k > g.NZ) return;
    const int jin = side ? g.NY : 2, jg = side ? g.NYY : 1;
    const long in = ix(g, i, jin, k), gh = ix(g, i, jg, k);
    if (i + gstart > heated_lo && i + gstart < heated_hi) {
        const double U1 = U.v[0][in], Jin = J[in];
        double rho = U1 * Jin;
        const double u = U.v[1][in] / U1;
        const double v = U.v[2][in] / U1;
        const double w = U.v[3][in] / U1;
        const double VV = u * u + v * v + w * w;
        const double P = (U.v[4][in] * Jin - 0.5 * rho * VV) * (K - 1);
        const double T = side ? T_high : T_low;
        rho = P / R / T;
        const double Jg = J[gh];
        U.v[0][gh] = rho / Jg;
        U.v[1][gh] = -rho * u / Jg;
        U.v[2][gh] = -rho * v / Jg;
        U.v[3][gh] = -rho * w / Jg;
        U.v[4][gh] = (P / (K - 1) + 0.5 * rho * VV) / Jg;
    } else {
#pragma unroll
        for (int q = 0; q < 5; q++) U.v[q][gh] = U.v[q][in];
    }
}

void launch_bc_y(const Geo &g, const Phys &p, int gstart, F5 U, const double *J, double T_low, double T_high,
                 int heated_lo, int heated_hi, hipStream_t s)
{
    dim3 b(64, 2);
    dim3 gr((g.NZ - 1 + b.x - 1) / b.x, 1, g.nx);
    hipLaunchKernelGGL(k_bc_y, gr, b, 0, s, g, p.K, p.R, gstart, U, J, T_low, T_high, heated_lo, heated_hi);
}

__global__ void k_bc_z(Geo g, int istart, F5 U)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;   // 0..NYYY
    const int i = istart + blockIdx.z;
    if (j > g.NYY + 1) return;
    const long r = ix(g, i, j, 0);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        U.v[q][r + 1] = U.v[q][r + g.NZ];
        U.v[q][r + g.NZZ] = U.v[q][r + 2];
    }
}

void launch_bc_z(const Geo &g, F5 U, hipStream_t s)
{
    const int istart = g.first ? 2 : 0;
    const int iend = g.last ? g.gend + 1 : g.gend + 3;
    dim3 b(64);
    dim3 gr((g.NYY + 2 + 63) / 64, 1, iend - istart + 1);
    hipLaunchKernelGGL(k_bc_z, gr, b, 0, s, g, istart, U);
}

__global__ void k_bc_x_periodic(Geo g, F5 U)
{
    const long plane = g.Ym * g.Zm;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int gpl = blockIdx.y;   // 0..2
    if (t >= plane) return;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        double *u = U.v[q];
        u[(long)gpl * plane + t] = u[(long)(g.gend - 2 + gpl) * plane + t];
        u[(long)(g.gend + 1 + gpl) * plane + t] = u[(long)(3 + gpl) * plane + t];
    }
}

void launch_bc_x_periodic(const Geo &g, F5 U, hipStream_t s)
{
    const long plane = g.Ym * g.Zm;
    dim3 b(256);
    dim3 gr((unsigned)((plane + 255) / 256), 3);
    hipLaunchKernelGGL(k_bc_x_periodic, gr, b, 0, s, g, U);
}

__global__ void k_replicate(double *dst, long plane, int nplanes)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    const double v = dst[t];
    for (int p = 1 + blockIdx.y; p < nplanes; p += gridDim.y) dst[(long)p * plane + t] = v;
}

void launch_replicate_planes(double *dst, long plane, int nplanes, hipStream_t s)
{
    dim3 gr((unsigned)((plane + 255) / 256), 16);
    hipLaunchKernelGGL(k_replicate, gr, dim3(256), 0, s, dst, plane, nplanes);
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

void launch_copy_field(double *dst, const double *src, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
}

// STREAM-style copy for the measured HBM bandwidth: 16 bytes per lane and
// iteration, grid-stride
__global__ __launch_bounds__(256) void k_stream_copy(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                                     long n2)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += stride) dst[t] = src[t];
}

void launch_stream_copy(double *dst, const double *src, long n, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, s, (double2 *)dst,
                       (const double2 *)src, n / 2);
}

// One block of the restart file: the interior cells of one field in (i, j, k)
// order with k fastest, Q = U*J (Output.cpp:183-201).
__global__ void k_pack_q(Geo g, const double *__restrict__ U, const double *__restrict__ J,
                         double *__restrict__ out)
{
    const long n = (long)g.nx * g.ny * g.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % g.nz);
    const long r = t / g.nz;
    const int j = (int)(r % g.ny), i = (int)(r / g.ny);
    const long c = ix(g, i + 3, j + 2, k + 2);
    out[t] = U[c] * J[c];
}

// The inverse (Initial_condition.cpp:196-212): U = Q/J, and U2 = U where the
// reference also starts the new level from the read one (U_ = U).
__global__ void k_unpack_q(Geo g, const double *__restrict__ in, const double *__restrict__ J,
                           double *__restrict__ U, double *__restrict__ U2)
{
    const long n = (long)g.nx * g.ny * g.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % g.nz);
    const long r = t / g.nz;
    const int j = (int)(r % g.ny), i = (int)(r / g.ny);
    const long c = ix(g, i + 3, j + 2, k + 2);
    const double v = in[t] / J[c];
    U[c] = v;
    if (U2) U2[c] = v;
}

void launch_pack_q(const Geo &g, const double *U, const double *J, double *out, hipStream_t s)
{
    const long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_pack_q, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, U, J, out);
}

void launch_unpack_q(const Geo &g, const double *in, const double *J, double *U, double *U2, hipStream_t s)
{
    const long n = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_unpack_q, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, in, J, U, U2);
}

}  // namespace cgli
