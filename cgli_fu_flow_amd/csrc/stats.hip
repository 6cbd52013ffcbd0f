// stats.hip -- wall-normal profiles of the channel on gfx950.
//
// The quantities Statistic.cpp:130-164 accumulates (U, V, W, T and the raw
// second moments UU, VV, WW, TT, UV, UT, VT) averaged over x and z: one
// workgroup per j, each thread accumulating a fixed stride of (i,k) in a
// fixed order, then a fixed-shape tree in LDS, so the result is the same
// bits on every run.  FP64 in every precision mode.
#include "cgli_internal.hpp"

namespace cgli {

constexpr int NPROF = 11;

__global__ void __launch_bounds__(256) k_profiles(Geo g, double K, double R, C5 U, const double *__restrict__ J,
                                                  double *__restrict__ out)
{
    __shared__ double sh[NPROF][256];
    const int j = 2 + blockIdx.x;
    const int t = threadIdx.x;
    const int n = g.nx * (g.NZ - 1);
    double s[NPROF];
#pragma unroll
    for (int a = 0; a < NPROF; a++) s[a] = 0.0;
    for (int q = t; q < n; q += 256) {
        const int i = 3 + q / (g.NZ - 1), k = 2 + q % (g.NZ - 1);
        const long c = ix(g, i, j, k);
        const double rho = U.v[0][c];
        const double u = U.v[1][c] / rho, v = U.v[2][c] / rho, w = U.v[3][c] / rho;
        const double VV = u * u + v * v + w * w;
        const double P = (U.v[4][c] - 0.5 * rho * VV) * (K - 1) * J[c];
        const double T = P / (rho * R) / J[c];
        s[0] += u; s[1] += v; s[2] += w; s[3] += T;
        s[4] += u * u; s[5] += v * v; s[6] += w * w; s[7] += T * T;
        s[8] += u * v; s[9] += u * T; s[10] += v * T;
    }
#pragma unroll
    for (int a = 0; a < NPROF; a++) sh[a][t] = s[a];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int a = 0; a < NPROF; a++) sh[a][t] += sh[a][t + h];
        }
        __syncthreads();
    }
    if (t < NPROF) {
        const int ny = g.NY - 1;
        out[t * ny + (j - 2)] = sh[t][0] * (1.0 / ((double)g.nx * (double)(g.NZ - 1)));
    }
}

void launch_profiles(const Geo &g, const Phys &p, C5 U, const double *J, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_profiles, dim3(g.NY - 1), dim3(256), 0, s, g, p.K, p.R, U, J, out);
}

// Per-block sums of f_q^2 over the interior cells (one block per (i, j)
// row of k), then a one-thread sum over blocks in index order.
__global__ void __launch_bounds__(256) k_sumsq5_rows(Geo g, C5 f, double *__restrict__ partial)
{
    __shared__ double sh[5][256];
    const int j = 2 + blockIdx.x % (g.NY - 1), i = 3 + blockIdx.x / (g.NY - 1);
    double s[5] = {0, 0, 0, 0, 0};
    for (int k = 2 + threadIdx.x; k <= g.NZ; k += 256) {
        const long c = ix(g, i, j, k);
#pragma unroll
        for (int q = 0; q < 5; q++) s[q] += f.v[q][c] * f.v[q][c];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[(long)blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

// one block: thread t sums blocks t, t+256, ... in order, then a fixed tree
__global__ void __launch_bounds__(256) k_sumsq5_final(const double *__restrict__ partial, long nblk,
                                                      double *__restrict__ out)
{
    __shared__ double sh[5][256];
    double s[5] = {0, 0, 0, 0, 0};
    for (long b = threadIdx.x; b < nblk; b += 256)
#pragma unroll
        for (int q = 0; q < 5; q++) s[q] += partial[b * 5 + q];
#pragma unroll
    for (int q = 0; q < 5; q++) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 5) out[threadIdx.x] = sh[threadIdx.x][0];
}

// partials: 5 x nx x ny doubles
void launch_sumsq5(const Geo &g, C5 f, double *partials, double *out_dev, hipStream_t s)
{
    const long nblk = (long)g.nx * (g.NY - 1);
    hipLaunchKernelGGL(k_sumsq5_rows, dim3((unsigned)nblk), dim3(256), 0, s, g, f, partials);
    hipLaunchKernelGGL(k_sumsq5_final, dim3(1), dim3(256), 0, s, partials, nblk, out_dev);
}

}  // namespace cgli
