// cnn.hip -- kernels of the learned initial guess: a 3-D convolutional
// network (15 -> 32 -> 32 -> 32 -> 5 channels, 3x3x3 kernels, GELU between
// layers) in BF16 with FP32 accumulation on the gfx950 matrix cores, its
// training kernels (input and weight gradients, masked MSE, Adam), and the
// snapshot ring it learns from.
//
// Activations are channel-last, [voxel][C] with C padded to 16 or 32 and
// voxel = ((b*D + d)*H + h)*W + w, W a multiple of 16.  Weights are
// [Cout_pad][27 taps][Cin_pad], tap = (dd+1)*9 + (dh+1)*3 + (dw+1) for the
// offsets dd, dh, dw in -1..1 (cross-correlation, as torch's conv3d with
// padding 1).  Outside a sample the input is zero, or, for the whole-field
// pass, periodic in D and W (x and z of the channel) and zero in H (y).
//
// A convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16:
//   C[co][n] += sum_k A[co][k] B[k][n],  k = (tap, ci),  n = 16 voxels of a row.
// Lane l holds row l%16 of A and column l%16 of B over the k values
// 8*(l/16)..+7 (16 contiguous bytes of one weight row, of one voxel's
// channels); the sum over k does not depend on how k is split across lanes as
// long as A and B use the same split.  The result holds rows 4*(l/16)..+3 of
// column l%16.
//   forward          z = conv(x, W) + b (FP32), a = GELU(z) rounded to BF16
//   input gradient   dx = conv(dz, W^T),  W^T[ci][26-tap][co] = W[co][tap][ci]
//   weight gradient  dW[co][tap][ci] = sum_v dz[v][co] x[v + off(tap)][ci]:
//                    per tap a GEMM over voxels with K = voxels, partial sums
//                    per block, then a fixed-order FP64 reduction, so the
//                    bits do not depend on scheduling
#include <algorithm>

#include "cgli_internal.hpp"

namespace cgli {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

// FP32 -> BF16 bits, round to nearest even (a NaN stays a quiet NaN)
__device__ __forceinline__ u16 f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (u16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// GELU of the convolution epilogue, z Phi(z), branch-free.  Phi comes from
// erfc(|z| / sqrt 2) by Abramowitz and Stegun 7.1.26,
//   erfc(x) = t (a1 + a2 t + ... + a5 t^4) exp(-x^2),  t = 1 / (1 + p x),
// absolute error <= 1.5e-7 -- within a few FP32 ulps of erff near 1 and far
// below the BF16 rounding its result goes through next.  Here p carries the
// 1 / sqrt 2 and the a_k the factor 1/2 of Phi = erfc(-z / sqrt 2) / 2.
__device__ __forceinline__ float gelu_fast(float z)
{
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f * 0.70710678118654752f, fabsf(z), 1.0f));   // v_rcp_f32
    const float poly =
        t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 0.5307027145f, -0.7265760135f), 0.7107068705f), -0.142248368f),
                 0.127414796f);
    const float h = poly * __builtin_amdgcn_exp2f((-0.72134752044448170f * z) * z);   // erfc(|z|/sqrt2)/2; v_exp_f32
    return z * (z >= 0.0f ? 1.0f - h : h);
}

// two FP32 values to packed BF16 (round to nearest even, the gfx950
// conversion instruction)
__device__ __forceinline__ unsigned pack_bf16x2(float a, float b)
{
    return (unsigned)__builtin_bit_cast(u16, (__bf16)a) | ((unsigned)__builtin_bit_cast(u16, (__bf16)b) << 16);
}
__device__ __forceinline__ float gelu_grad(float z)
{
    const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
    return Phi + z * phi;
}

// Per-lane K fragment of one tap: CIN = 32 -> 8 channels (16 bytes) for
// v_mfma_f32_16x16x32_bf16; CIN = 16 -> 4 channels (8 bytes) for
// v_mfma_f32_16x16x16_bf16.  Fragments travel as 32-bit words.
typedef short s16x4 __attribute__((ext_vector_type(4)));
template <int CIN> struct KF;
template <> struct KF<32> {
    typedef uint4 raw;
    static __device__ __forceinline__ raw zero() { return make_uint4(0u, 0u, 0u, 0u); }
    static __device__ __forceinline__ raw sel(bool ok, raw v)
    {
        return make_uint4(ok ? v.x : 0u, ok ? v.y : 0u, ok ? v.z : 0u, ok ? v.w : 0u);
    }
    static __device__ __forceinline__ f32x4 mfma(raw a, raw b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    }
};
template <> struct KF<16> {
    typedef uint2 raw;
    static __device__ __forceinline__ raw zero() { return make_uint2(0u, 0u); }
    static __device__ __forceinline__ raw sel(bool ok, raw v) { return make_uint2(ok ? v.x : 0u, ok ? v.y : 0u); }
    static __device__ __forceinline__ f32x4 mfma(raw a, raw b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4, a), __builtin_bit_cast(s16x4, b), c,
                                                         0, 0, 0);
    }
};

// Block tile of a convolution: CT_D planes x CT_H rows x CT_W voxels of the
// output, computed by 8 waves of 1 plane x 4 rows x 16 voxels each.  Its
// input, the tile plus a 1-voxel halo (CS_D planes of CS_H x CS_W voxels),
// sits in LDS in CS_D plane slots.
constexpr int CT_D = 2, CT_H = 8, CT_W = 32, CT_THREADS = 512;
constexpr int CS_D = CT_D + 2, CS_H = CT_H + 2, CS_W = CT_W + 2;
constexpr int CS_P = CS_H * CS_W;   // voxels of a staged plane (340)
constexpr int CS_N = CS_D * CS_P;   // 1360

struct ConvTile { int b, d0, h0, w0; };

// tiles are numbered plane-fastest: tile T and T + 1 of one column (same
// b, h0, w0) are CT_D planes apart
__device__ __forceinline__ ConvTile conv_tile(int T, const NetDims &m)
{
    const int nd = (m.D + CT_D - 1) / CT_D, nh = (m.H + CT_H - 1) / CT_H, nw = (m.W + CT_W - 1) / CT_W;
    ConvTile c;
    c.d0 = (T % nd) * CT_D;
    int r = T / nd;
    c.h0 = (r % nh) * CT_H;
    r /= nh;
    c.w0 = (r % nw) * CT_W;
    c.b = r / nw;
    return c;
}

// 16 bytes (8 channels) into a [kq][item][CPL] LDS array: one fragment when
// CPL = 8, two (kq and kq + 1, `stride` elements apart) when CPL = 4
template <int CPL>
__device__ __forceinline__ void put_chunk(u16 *dst, int stride, uint4 v)
{
    if constexpr (CPL == 8) {
        *reinterpret_cast<uint4 *>(dst) = v;
    } else {
        *reinterpret_cast<uint2 *>(dst) = make_uint2(v.x, v.y);
        *reinterpret_cast<uint2 *>(dst + stride) = make_uint2(v.z, v.w);
    }
}

// LDS stride (elements) of one kq quarter of the staged input: padded so
// that consecutive quarters start 64 bytes apart modulo the 256 bytes the
// banks span, which keeps the staging stores free of bank conflicts
template <int CIN>
constexpr int stage_plane()
{
    constexpr int b = CS_N * (CIN / 4) * 2;
    return (b + ((64 - b % 256) + 256) % 256) / 2;
}

// 16 zero bytes: the source of staged chunks outside the sample
__device__ const uint4 k_zero16 = {0u, 0u, 0u, 0u};

// Staging moves two planes at a time, CT_THREADS threads x PF 16-byte chunks.
// Chunk e = thread + CT_THREADS * s is chunk e % CPR of staged row e / CPR
// (CPR chunks per row of CS_W voxels): plane pp = row / CS_H (0 or 1 of the
// pair), row r = row % CS_H, voxel x of the row and 8-channel group ch.
// code[s] = pp << 24 | r << 16 | x << 8 | ch (-1 past the last chunk) and
// soff[s], its LDS element offset from the pair's first slot, do not depend
// on the tile.
template <int CIN, int PF>
__device__ __forceinline__ void stage_codes(int *code, int *soff)
{
    constexpr int CPL = CIN / 4, CPV = CIN / 8, CPR = CS_W * CPV, NCH = 2 * CS_H * CPR, KS = stage_plane<CIN>();
#pragma unroll
    for (int s = 0; s < PF; s++) {
        const int e = threadIdx.x + CT_THREADS * s;
        const int R = e / CPR, j = e % CPR, pp = R / CS_H, r = R % CS_H, x = j / CPV, ch = j % CPV;
        code[s] = e < NCH ? pp << 24 | r << 16 | x << 8 | ch : -1;
        soff[s] = (8 * ch / CPL) * KS + (pp * CS_P + r * CS_W + x) * CPL;
    }
}

// the chunks of tile-relative planes p0 and p0 + 1 of tile c (plane p holds
// d0 - 1 + p) into registers: zero outside the sample unless PER (periodic
// in d and w), always zero outside in h
template <int CIN, int PF, bool PER>
__device__ __forceinline__ void stage_load(uint4 *pf, const u16 *__restrict__ X, const NetDims &m, const ConvTile &c,
                                           int p0, const int *code)
{
    int dA = c.d0 - 1 + p0, dB = dA + 1;
    if constexpr (PER) {
        dA = dA < 0 ? dA + m.D : (dA >= m.D ? dA - m.D : dA);
        dB = dB < 0 ? dB + m.D : (dB >= m.D ? dB - m.D : dB);
        if (m.D < 2) dA = dB = 0;
    }
    const bool okA = dA >= 0 && dA < m.D, okB = dB >= 0 && dB < m.D;
    const long plane = (long)m.H * m.W * CIN;
    const u16 *PA = X + ((long)c.b * m.D + (okA ? dA : 0)) * plane;
    const u16 *PB = X + ((long)c.b * m.D + (okB ? dB : 0)) * plane;
    const int hm = c.h0 - 1, wm = c.w0 - 1;
    const bool inner = c.w0 >= 1 && c.w0 + CS_W - 1 <= m.W;   // voxels w0-1 .. w0+CT_W inside, no wrap
#pragma unroll
    for (int s = 0; s < PF; s++) {
        const int cd = code[s];
        const int pp = (cd >> 24) & 1, h = hm + ((cd >> 16) & 255), ch = cd & 255;
        int w = wm + ((cd >> 8) & 255);
        bool ok = cd >= 0 && (pp ? okB : okA) && (unsigned)h < (unsigned)m.H;
        if constexpr (PER) {
            w = w < 0 ? w + m.W : w;
            w = w >= m.W ? w - m.W : w;
            if (m.W < CS_W) w %= m.W;   // narrow samples: more than one period in a staged row
        } else {
            ok = ok && (inner || (unsigned)w < (unsigned)m.W);
        }
        const u16 *src = (pp ? PB : PA) + (long)(h * m.W + w) * CIN + 8 * ch;
        pf[s] = *(ok ? reinterpret_cast<const uint4 *>(src) : &k_zero16);
    }
}

// the chunks of stage_load into LDS plane slots slot0 and slot0 + 1
template <int CIN, int PF>
__device__ __forceinline__ void stage_store(u16 *sx, const uint4 *pf, int slot0, const int *code, const int *soff)
{
    constexpr int CPL = CIN / 4, KS = stage_plane<CIN>();
    u16 *base = sx + slot0 * CS_P * CPL;
#pragma unroll
    for (int s = 0; s < PF; s++)
        if (code[s] >= 0) put_chunk<CPL>(base + soff[s], KS, pf[s]);
}

// y = conv(x, W) + bias for all voxels.  Persistent blocks of 8 waves, one
// per CU: the layer's weights (COUT x 27 x CIN BF16) are staged once per
// block in LDS as [tap][kq][co][CPL], and the staged input as
// [kq][voxel][CPL] (both put a wave's 16 lanes of one kq on consecutive
// fragments).  Each block walks a contiguous range of tiles, so that
// consecutive tiles mostly continue one column: the two planes they share
// stay in their LDS slots (a ring of four) and only two new planes are
// staged, loaded into registers while the previous tile is computed.  A
// wave's K loop runs over the 9 (dd, dw) tap pairs: the 3 taps dh = -1..1
// of the pair are held in registers and each of the 6 input rows its 4
// output rows touch is read once from LDS and used by up to 3 of them.
// Blocks on one XCD take neighbouring ranges (workgroups are dealt to the 8
// XCDs round robin), so that halo rows they share meet in that XCD's L2.
// Outputs (either may be null): Z FP32 [v][COUT] (the pre-activation),
// A BF16 [v][COUT] = GELU(Z) when ACT, else Z; channels co >= cout are
// written as 0.
template <int CIN, int MT, bool ACT, bool PER>
__global__ void __launch_bounds__(CT_THREADS) k_conv(const u16 *__restrict__ X, const u16 *__restrict__ Wt,
                                                     const float *__restrict__ bias, float *__restrict__ Z,
                                                     u16 *__restrict__ A, NetDims m, int cout, int ntiles)
{
    constexpr int COUT = 16 * MT;
    constexpr int CPL = CIN / 4;                                           // channels per lane fragment
    constexpr int PF = (2 * CS_H * CS_W * CIN / 8 + CT_THREADS - 1) / CT_THREADS;   // chunks per thread, 2 planes
    constexpr int KS = stage_plane<CIN>();                                 // LDS stride of a kq quarter
    typedef KF<CIN> K;
    typedef typename K::raw R;
    __shared__ __align__(16) u16 sw[27 * CIN * COUT];
    __shared__ __align__(16) u16 sx[4 * KS];
    for (int t = threadIdx.x; t < COUT * 27 * CIN / 8; t += CT_THREADS) {
        const int e = 8 * t, co = e / (27 * CIN), r = e % (27 * CIN), tap = r / CIN, kq = (r % CIN) / CPL;
        put_chunk<CPL>(sw + ((tap * 4 + kq) * COUT + co) * CPL, COUT * CPL, reinterpret_cast<const uint4 *>(Wt)[t]);
    }
    const int G = gridDim.x;
    const int rix = G % 8 == 0 ? (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const int T0 = (int)((long)rix * ntiles / G), T1 = (int)((long)(rix + 1) * ntiles / G);
    const int nd = (m.D + CT_D - 1) / CT_D;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = lane & 15, kq = lane >> 4;
    const int pd = wv & 1, hg = (wv >> 1) & 1, wh = wv >> 2;
    const u16 *xq = sx + kq * KS;
    const u16 *wq = sw + (kq * COUT + n) * CPL;
    float bz[MT][4];   // this lane's biases (0 for padded channels)
#pragma unroll
    for (int u = 0; u < MT; u++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int ch = 16 * u + 4 * kq + rr;
            bz[u][rr] = bias && ch < cout ? bias[ch] : 0.0f;
        }
    int code[PF], soff[PF];
    stage_codes<CIN, PF>(code, soff);
    uint4 pf[PF];
    int base = 0;   // LDS slot of tile-relative plane 0
    if (T0 < T1) stage_load<CIN, PF, PER>(pf, X, m, conv_tile(T0, m), 0, code);
    for (int T = T0; T < T1; T++) {
        const ConvTile c = conv_tile(T, m);
        __syncthreads();   // the previous tile's LDS reads (and the weights' staging) are done
        if (T > T0 && T % nd != 0) {
            // same column as tile T - 1: its planes 2, 3 are this tile's 0, 1
            base = (base + 2) & 3;
            stage_store<CIN, PF>(sx, pf, (base + 2) & 3, code, soff);
        } else {
            base = 0;
            stage_store<CIN, PF>(sx, pf, 0, code, soff);
            stage_load<CIN, PF, PER>(pf, X, m, c, 2, code);
            stage_store<CIN, PF>(sx, pf, 2, code, soff);
        }
        __syncthreads();
        if (T + 1 < T1) stage_load<CIN, PF, PER>(pf, X, m, conv_tile(T + 1, m), (T + 1) % nd != 0 ? 2 : 0, code);
        const int d = c.d0 + pd, hb = c.h0 + 4 * hg, w = c.w0 + 16 * wh + n;
        if (d >= m.D || hb >= m.H || c.w0 + 16 * wh >= m.W) continue;   // wave-uniform
        f32x4 acc[4][MT];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int u = 0; u < MT; u++) acc[q][u] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int dd = 0; dd < 3; dd++) {   // not unrolled: bounds the LDS reads in flight
            const u16 *xp = xq + (((pd + dd + base) & 3) * CS_P + 4 * hg * CS_W + 16 * wh + n) * CPL;
#pragma unroll
            for (int dw = 0; dw < 3; dw++) {
                R a[3][MT];
#pragma unroll
                for (int dh = 0; dh < 3; dh++)
#pragma unroll
                    for (int u = 0; u < MT; u++)
                        a[dh][u] = *reinterpret_cast<const R *>(wq + ((dd * 9 + dh * 3 + dw) * 4 * COUT + 16 * u) * CPL);
#pragma unroll
                for (int rp = 0; rp < 6; rp++) {
                    const R bq = *reinterpret_cast<const R *>(xp + (rp * CS_W + dw) * CPL);
#pragma unroll
                    for (int dh = 0; dh < 3; dh++) {
                        const int q = rp - dh;
                        if (q < 0 || q >= 4) continue;
#pragma unroll
                        for (int u = 0; u < MT; u++) acc[q][u] = K::mfma(a[dh][u], bq, acc[q][u]);
                    }
                }
            }
        }
        const bool padded = cout < COUT;   // uniform
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (hb + q >= m.H) break;
            const long v = (((long)c.b * m.D + d) * m.H + hb + q) * m.W + w;
#pragma unroll
            for (int u = 0; u < MT; u++) {
                const int cb = 16 * u + 4 * kq;   // channels cb .. cb+3 of this lane
                float z[4], g[4];
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    z[rr] = acc[q][u][rr] + bz[u][rr];
                    g[rr] = ACT ? gelu_fast(z[rr]) : z[rr];
                }
                if (padded) {
#pragma unroll
                    for (int rr = 0; rr < 4; rr++)
                        if (cb + rr >= cout) z[rr] = g[rr] = 0.0f;
                }
                if (Z) *reinterpret_cast<float4 *>(Z + v * COUT + cb) = make_float4(z[0], z[1], z[2], z[3]);
                if (A)
                    *reinterpret_cast<uint2 *>(A + v * COUT + cb) = make_uint2(pack_bf16x2(g[0], g[1]), pack_bf16x2(g[2], g[3]));
            }
        }
    }
}

// dz = da * GELU'(z): BF16 for the next convolutions, FP32 for the bias
__global__ void k_gelu_back(const float *__restrict__ dA, const float *__restrict__ Z, u16 *__restrict__ dZ,
                            float *__restrict__ dZf, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float g = dA[t] * gelu_grad(Z[t]);
    dZ[t] = f2bf(g);
    dZf[t] = g;
}

__global__ void k_to_bf16(const float *__restrict__ x, u16 *__restrict__ y, long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = f2bf(x[t]);
}

// W^T[ci][26-tap][co] = W[co][tap][ci]   (Cin_pad x 27 x Cout_pad)
__global__ void k_transpose_w(const u16 *__restrict__ W, u16 *__restrict__ WT, int cin_pad, int cout_pad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cin_pad * 27 * cout_pad) return;
    const int co = t % cout_pad, tap = (t / cout_pad) % 27, ci = t / (cout_pad * 27);
    WT[t] = W[((long)co * 27 + (26 - tap)) * cin_pad + ci];
}

// Weight-gradient partials.  Block (bx, tap) sums dz[v][co] x[v + off(tap)][ci]
// over its voxel range [bx*per, (bx+1)*per) in chunks of CH voxels, staged
// transposed in LDS ([channel][voxel]) so that a lane's 8 consecutive voxels
// of one channel are 16 contiguous bytes.  One wave per 16x16 (co, ci) tile.
// part[bx][tap][co][ci].
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(const u16 *__restrict__ dZ, const u16 *__restrict__ X, NetDims m,
                                               long nvox, long per, float *__restrict__ part)
{
    constexpr int CH = 128, LD = CH + 8;
    constexpr int MT = COUT / 16, NT = CIN / 16, NW = MT * NT;
    constexpr int VZ = COUT / 8, VX = CIN / 8;    // 16-byte vectors per voxel
    __shared__ __align__(16) u16 sdz[COUT * LD];
    __shared__ __align__(16) u16 sx[CIN * LD];
    __shared__ long src[CH];
    const int tap = blockIdx.y;
    const int dd = tap / 9 - 1, dh = (tap / 3) % 3 - 1, dw = tap % 3 - 1;
    const long v0 = (long)blockIdx.x * per;
    const long v1 = v0 + per < nvox ? v0 + per : nvox;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int mt = wave / NT, nt = wave % NT;
    const int n = lane & 15, kq = lane >> 4;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long vb = v0; vb < v1; vb += CH) {
        __syncthreads();
        if (threadIdx.x < CH) {
            const long v = vb + threadIdx.x;
            long sidx = -1;
            if (v < v1) {
                const int w = (int)(v % m.W), h = (int)((v / m.W) % m.H);
                const int d = (int)((v / ((long)m.W * m.H)) % m.D), b = (int)(v / ((long)m.W * m.H * m.D));
                const int d2 = d + dd, h2 = h + dh, w2 = w + dw;
                if (d2 >= 0 && d2 < m.D && h2 >= 0 && h2 < m.H && w2 >= 0 && w2 < m.W)
                    sidx = (((long)b * m.D + d2) * m.H + h2) * m.W + w2;
            }
            src[threadIdx.x] = sidx;
        }
        __syncthreads();
        // BF16 data moves as 32-bit words (two channels each): element
        // access to __bf16 vectors is avoided
        for (int t = threadIdx.x; t < CH * (VZ + VX); t += blockDim.x) {
            const int vv = t % CH, q = t / CH;
            const long v = vb + vv;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            const u16 *srcp = nullptr;
            u16 *dst;
            if (q < VZ) {
                if (v < v1) srcp = dZ + v * COUT + 8 * q;
                dst = sdz + 8 * q * LD + vv;
            } else {
                const long s = src[vv];
                if (s >= 0) srcp = X + s * CIN + 8 * (q - VZ);
                dst = sx + 8 * (q - VZ) * LD + vv;
            }
            if (srcp) val = *reinterpret_cast<const uint4 *>(srcp);
            const unsigned wd[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
            for (int e = 0; e < 8; e++) dst[e * LD] = (u16)(wd[e >> 1] >> (16 * (e & 1)));
        }
        __syncthreads();
        if (wave < NW) {
#pragma unroll
            for (int kc = 0; kc < CH / 32; kc++) {
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sdz[(16 * mt + n) * LD + 32 * kc + 8 * kq]);
                const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(&sx[(16 * nt + n) * LD + 32 * kc + 8 * kq]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
            }
        }
    }
    if (wave < NW) {
        float *o = part + ((long)blockIdx.x * 27 + tap) * COUT * CIN;
#pragma unroll
        for (int r = 0; r < 4; r++) o[(16 * mt + 4 * kq + r) * CIN + 16 * nt + n] = acc[r];
    }
}

// dW[co][tap][ci] = sum over blocks, in block order; padded entries are 0
__global__ void k_wgrad_reduce(const float *__restrict__ part, int nblk, int cout_pad, int cin_pad, int cout,
                               int cin, float *__restrict__ dW)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cout_pad * 27 * cin_pad) return;
    const int ci = t % cin_pad, tap = (t / cin_pad) % 27, co = t / (cin_pad * 27);
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(((long)b * 27 + tap) * cout_pad + co) * cin_pad + ci];
    dW[t] = (co < cout && ci < cin) ? (float)s : 0.0f;
}

// db[co] = sum_v dz[v][co] (dz FP32 [v][CP]) in two fixed-order stages:
// block b sums its voxel range per channel (thread t: channel t % CP,
// voxels strided by 256 / CP, then a tree over those lanes), then one block
// sums the blocks in index order.
template <int CP>
__global__ void __launch_bounds__(256) k_bias_part(const float *__restrict__ dZ, long nvox, long per,
                                                   double *__restrict__ part)
{
    constexpr int VL = 256 / CP;
    __shared__ double sh[256];
    const int c = threadIdx.x % CP, vl = threadIdx.x / CP;
    const long v0 = (long)blockIdx.x * per, v1 = v0 + per < nvox ? v0 + per : nvox;
    double s = 0.0;
    for (long v = v0 + vl; v < v1; v += VL) s += dZ[v * CP + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = VL / 2; h > 0; h >>= 1) {
        if (vl < h) sh[threadIdx.x] += sh[threadIdx.x + h * CP];
        __syncthreads();
    }
    if (vl == 0) part[(long)blockIdx.x * CP + c] = sh[threadIdx.x];
}

__global__ void k_bias_final(const double *__restrict__ part, int nblk, int cp, int cout, float *__restrict__ db)
{
    const int c = threadIdx.x;
    if (c >= cp) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(long)b * cp + c];
    db[c] = c < cout ? (float)s : 0.0f;
}

// Masked mean-square error over the first 5 of 16 output channels:
// partial[block] = sum of e^2, dY = 2 e inv on masked voxels (0 elsewhere),
// inv = 1 / (5 x masked voxels)
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             const unsigned char *__restrict__ mask, long nvox, float inv,
                                             float *__restrict__ dY, double *__restrict__ partial)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < nvox * 16; t += (long)gridDim.x * 256) {
        const int c = (int)(t & 15);
        const long v = t >> 4;
        float g = 0.0f;
        if (c < 5 && mask[v]) {
            const float e = Y[t] - T[v * 5 + c];
            s += (double)e * e;
            g = 2.0f * e * inv;
        }
        dY[t] = g;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out = scale x sum of partial[0..n) in index order (one thread)
__global__ void k_sum_partials(const double *__restrict__ partial, int n, double scale, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < n; b++) s += partial[b];
        *out = s * scale;
    }
}

// Adam on the FP32 master parameters; refreshes their BF16 copy
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m1, float *__restrict__ m2,
                       const float *__restrict__ g, u16 *__restrict__ wb, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.0f - b1) * gt;
    const float b = b2 * m2[t] + (1.0f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = b;
    const float nw = w[t] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[t] = nw;
    wb[t] = f2bf(nw);
}

// ---------------------------------------------------------------- snapshots
// Ring slot = the physical-step increment per unit volume, dq = J (U - Uq),
// over the interior cells (U holds conserved variables over J), FP32,
// [cell][5] with cell = ((i-3)*ny + (j-2))*nz + (k-2).
__global__ void k_store(Geo g, C5 U, C5 Uq, const double *__restrict__ J, float *__restrict__ S)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
    const long v = ((long)(i - 3) * g.ny + (j - 2)) * g.nz + (k - 2);
    const double Jc = J[c];
#pragma unroll
    for (int q = 0; q < 5; q++) S[v * 5 + q] = (float)((U.v[q][c] - Uq.v[q][c]) * Jc);
}

// Network input and target of sample b.  org[b] = (t, i0, j0, k0): record t
// (or, when whole_rec >= 0, org = (whole_rec, 0, 0, 0) for the one sample)
// of the ring (slot t % cap), crop origin (i0, j0, k0), wrapping in i and k.
// rs[slot][15] holds each record's inverse scales 1/s1, 1/sd, 1/s2 (five
// each): the field RMS of dq_t, of d1_t = dq_t - dq_{t-1} and of
// d2_t = d1_t - d1_{t-1}, per variable.  With record t's scales:
//   x[0..4]   = dq_t / s1
//   x[5..9]   = d1_t / sd
//   x[10..14] = d2_t / s2,   x[15] = 0
//   target    = d2_{t+1} / s2 = ((dq_{t+1} - dq_t) - d1_t) / s2  (T not null)
// so that a zero output is the quadratic extrapolation dq_{t+1} = dq_t + d1_t.
// mask[v] (when not null): 1 where the voxel's `halo`-voxel neighbourhood
// lies in the crop or beyond a wall (j outside 0..ny-1), else 0.
__global__ void k_gather(const float *__restrict__ S, const float *__restrict__ rs, int cap, long vol, int nx,
                         int ny, int nz, const int *__restrict__ org, int whole_rec, NetDims m, int halo,
                         u16 *__restrict__ X, float *__restrict__ T, unsigned char *__restrict__ mask)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nv = (long)m.B * m.D * m.H * m.W;
    if (t >= nv) return;
    long cell;
    int rec, d = 0, h = 0, w = 0, oj = 0;
    if (whole_rec >= 0) {
        // one sample, the whole grid of record whole_rec: voxel t is cell t
        cell = t;
        rec = whole_rec;
    } else {
        // crops (a batch of crops holds fewer than 2^31 voxels)
        const int tt = (int)t;
        w = tt % m.W;
        int r = tt / m.W;
        h = r % m.H;
        r /= m.H;
        d = r % m.D;
        const int b = r / m.D;
        rec = org[4 * b];
        oj = org[4 * b + 2];
        const int i = (org[4 * b + 1] + d) % nx, k = (org[4 * b + 3] + w) % nz;
        cell = ((long)i * ny + oj + h) * nz + k;
    }
    const float *sc = rs + (long)(rec % cap) * NET_NSCALE;
    const float *s0 = S + ((long)(rec % cap) * vol + cell) * 5;
    const float *s1 = S + ((long)((rec - 1) % cap) * vol + cell) * 5;
    const float *s2 = S + ((long)((rec - 2) % cap) * vol + cell) * 5;
    float f[16];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const float a = s0[q], bq = s1[q], c = s2[q];
        const float d1 = a - bq, d1p = bq - c;
        f[q] = a * sc[q];
        f[5 + q] = d1 * sc[5 + q];
        f[10 + q] = (d1 - d1p) * sc[10 + q];
    }
    f[15] = 0.0f;
    uint4 *x = reinterpret_cast<uint4 *>(X + t * 16);
    x[0] = make_uint4(pack_bf16x2(f[0], f[1]), pack_bf16x2(f[2], f[3]), pack_bf16x2(f[4], f[5]), pack_bf16x2(f[6], f[7]));
    x[1] = make_uint4(pack_bf16x2(f[8], f[9]), pack_bf16x2(f[10], f[11]), pack_bf16x2(f[12], f[13]),
                      pack_bf16x2(f[14], f[15]));
    if (T) {
        const float *s3 = S + ((long)((rec + 1) % cap) * vol + cell) * 5;
#pragma unroll
        for (int q = 0; q < 5; q++) T[t * 5 + q] = ((s3[q] - s0[q]) - (s0[q] - s1[q])) * sc[10 + q];
    }
    if (mask) {
        const bool okd = d >= halo && d < m.D - halo;
        const bool okw = w >= halo && w < m.W - halo;
        const bool okh = (h >= halo || oj == 0) && (h < m.H - halo || oj + m.H == ny);
        mask[t] = (okd && okw && okh) ? 1 : 0;
    }
}

// Host-supplied FP32 input [v][15] -> BF16 [v][16]
__global__ void k_pack_input(const float *__restrict__ x, u16 *__restrict__ X, long nv)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
#pragma unroll
    for (int q = 0; q < 15; q++) X[t * 16 + q] = f2bf(x[t * 15 + q]);
    X[t * 16 + 15] = 0;
}

// Per-block sums over vol cells of dq_a^2, d1^2 and d2^2 (five each) with
// d1 = dq_a - dq_b, d2 = d1 - (dq_b - dq_c), FP32 differences, FP64 sums
__global__ void __launch_bounds__(256) k_record_sums(const float *__restrict__ Sa, const float *__restrict__ Sb,
                                                     const float *__restrict__ Sc, long vol,
                                                     double *__restrict__ partial)
{
    __shared__ double sh[NET_NSCALE][256];
    double s[NET_NSCALE];
#pragma unroll
    for (int q = 0; q < NET_NSCALE; q++) s[q] = 0.0;
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < vol; v += (long)gridDim.x * 256)
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const float a = Sa[v * 5 + q], b = Sb[v * 5 + q], c = Sc[v * 5 + q];
            const float d1 = a - b, d2 = d1 - (b - c);
            s[q] += (double)a * a;
            s[5 + q] += (double)d1 * d1;
            s[10 + q] += (double)d2 * d2;
        }
#pragma unroll
    for (int q = 0; q < NET_NSCALE; q++) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < NET_NSCALE; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < NET_NSCALE) partial[blockIdx.x * NET_NSCALE + threadIdx.x] = sh[threadIdx.x][0];
}

// Inverse scales of a record from the block sums, in block order: 1/RMS, or
// 1 where an RMS is 0 or its earlier records do not exist (nprev < 1 for
// d1, < 2 for d2)
__global__ void k_record_scales(const double *__restrict__ partial, int nblk, long vol, int nprev,
                                float *__restrict__ rs)
{
    const int q = threadIdx.x;
    if (q >= NET_NSCALE) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[b * NET_NSCALE + q];
    const double rms = sqrt(s / (double)vol);
    rs[q] = (rms > 0.0 && nprev >= q / 5) ? (float)(1.0 / rms) : 1.0f;
}

// Initial guess U_ = U + dq (+ (dq - dq_prev) + s2 Y / J when prev and rs
// are given), with dq = U - Uq in FP64 and dq_prev = (the ring's previous
// record) / J: the linear extrapolation, or the quadratic one plus the
// network's correction of the second difference (Y [cell][16], first 5
// channels, per unit volume; rs = the latest record's inverse scales).
__global__ void k_guess(Geo g, C5 U, C5 Uq, const double *__restrict__ J, const float *__restrict__ prev,
                        const float *__restrict__ Y, const float *__restrict__ rs, F5 U_)
{
    const int k = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 3 + blockIdx.z;
    if (k > g.NZ || j > g.NY) return;
    const long c = ix(g, i, j, k);
    const long v = ((long)(i - 3) * g.ny + (j - 2)) * g.nz + (k - 2);
    const double Jc = J[c];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const double u = U.v[q][c];
        const double dq = u - Uq.v[q][c];
        double next = dq;
        if (prev) {
            next += dq - (double)prev[v * 5 + q] / Jc;
            if (rs) next += (double)Y[v * 16 + q] / ((double)rs[10 + q] * Jc);
        }
        U_.v[q][c] = u + next;
    }
}

// ---------------------------------------------------------------- launchers
static unsigned blocks_of(long n, int b) { return (unsigned)((n + b - 1) / b); }

// one persistent block per CU (a block's LDS holds the weights and one
// staged tile)
static int conv_grid(int ntiles)
{
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
    }
    return std::max(1, std::min(ntiles, ncu));
}

template <bool PER>
static int conv_launch(int cin_pad, int cout_pad, bool act, const u16 *x, const u16 *w, const float *bias, float *Z,
                       u16 *a, const NetDims &m, int cout, hipStream_t s)
{
    const long nt = (long)m.B * ((m.D + CT_D - 1) / CT_D) * ((m.H + CT_H - 1) / CT_H) * ((m.W + CT_W - 1) / CT_W);
    if (nt < 1 || nt > 0x7fffffffL) return -1;
    const int ntiles = (int)nt;
    const dim3 gr(conv_grid(ntiles)), bl(CT_THREADS);
    if (cin_pad == 16 && cout_pad == 32 && act)
        hipLaunchKernelGGL((k_conv<16, 2, true, PER>), gr, bl, 0, s, x, w, bias, Z, a, m, cout, ntiles);
    else if (cin_pad == 32 && cout_pad == 32 && act)
        hipLaunchKernelGGL((k_conv<32, 2, true, PER>), gr, bl, 0, s, x, w, bias, Z, a, m, cout, ntiles);
    else if (cin_pad == 32 && cout_pad == 16 && !act)
        hipLaunchKernelGGL((k_conv<32, 1, false, PER>), gr, bl, 0, s, x, w, bias, Z, a, m, cout, ntiles);
    else if (cin_pad == 16 && cout_pad == 32 && !act)
        hipLaunchKernelGGL((k_conv<16, 2, false, PER>), gr, bl, 0, s, x, w, bias, Z, a, m, cout, ntiles);
    else if (cin_pad == 32 && cout_pad == 32 && !act)
        hipLaunchKernelGGL((k_conv<32, 2, false, PER>), gr, bl, 0, s, x, w, bias, Z, a, m, cout, ntiles);
    else
        return -1;
    return 0;
}

int net_conv(int cin_pad, int cout_pad, bool act, const void *X, const void *W, const float *bias, float *Z, void *A,
             const NetDims &m, int cout, int per, hipStream_t s)
{
    if (m.W % 16) return -1;
    const u16 *x = static_cast<const u16 *>(X), *w = static_cast<const u16 *>(W);
    u16 *a = static_cast<u16 *>(A);
    return per ? conv_launch<true>(cin_pad, cout_pad, act, x, w, bias, Z, a, m, cout, s)
               : conv_launch<false>(cin_pad, cout_pad, act, x, w, bias, Z, a, m, cout, s);
}

void net_gelu_back(const float *dA, const float *Z, void *dZ, float *dZf, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gelu_back, dim3(blocks_of(n, 256)), dim3(256), 0, s, dA, Z, static_cast<u16 *>(dZ), dZf, n);
}

void net_to_bf16(const float *x, void *y, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_of(n, 256)), dim3(256), 0, s, x, static_cast<u16 *>(y), n);
}

void net_transpose_w(const void *W, void *WT, int cin_pad, int cout_pad, hipStream_t s)
{
    const long n = (long)cin_pad * 27 * cout_pad;
    hipLaunchKernelGGL(k_transpose_w, dim3(blocks_of(n, 256)), dim3(256), 0, s, static_cast<const u16 *>(W),
                       static_cast<u16 *>(WT), cin_pad, cout_pad);
}

int net_wgrad(int cin_pad, int cout_pad, const void *dZ, const void *X, const NetDims &m, int nblk, float *part,
              int cout, int cin, float *dW, hipStream_t s)
{
    const long nvox = (long)m.B * m.D * m.H * m.W;
    const long per = (nvox + nblk - 1) / nblk;
    const dim3 gr(nblk, 27), bl(256);
    const u16 *dz = static_cast<const u16 *>(dZ), *x = static_cast<const u16 *>(X);
    if (cin_pad == 16 && cout_pad == 32)
        hipLaunchKernelGGL((k_wgrad<16, 32>), gr, bl, 0, s, dz, x, m, nvox, per, part);
    else if (cin_pad == 32 && cout_pad == 32)
        hipLaunchKernelGGL((k_wgrad<32, 32>), gr, bl, 0, s, dz, x, m, nvox, per, part);
    else if (cin_pad == 32 && cout_pad == 16)
        hipLaunchKernelGGL((k_wgrad<32, 16>), gr, bl, 0, s, dz, x, m, nvox, per, part);
    else
        return -1;
    const long n = (long)cout_pad * 27 * cin_pad;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_of(n, 256)), dim3(256), 0, s, part, nblk, cout_pad, cin_pad, cout,
                       cin, dW);
    return 0;
}

// part: NET_BIAS_BLOCKS x cpad doubles
void net_bias_grad(const float *dZ, long nvox, int cpad, int cout, float *db, double *part, hipStream_t s)
{
    const int nblk = (int)std::max(1L, std::min<long>(NET_BIAS_BLOCKS, (nvox + 255) / 256));
    const long per = (nvox + nblk - 1) / nblk;
    if (cpad == 32)
        hipLaunchKernelGGL((k_bias_part<32>), dim3(nblk), dim3(256), 0, s, dZ, nvox, per, part);
    else
        hipLaunchKernelGGL((k_bias_part<16>), dim3(nblk), dim3(256), 0, s, dZ, nvox, per, part);
    hipLaunchKernelGGL(k_bias_final, dim3(1), dim3(64), 0, s, part, nblk, cpad, cout, db);
}

void net_mse(const float *Y, const float *T, const unsigned char *mask, long nvox, float inv, float *dY,
             double *partial, int nblk, double *loss, double loss_scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_mse, dim3(nblk), dim3(256), 0, s, Y, T, mask, nvox, inv, dY, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, nblk, loss_scale, loss);
}

void net_adam(float *w, float *m1, float *m2, const float *g, void *wb, long n, float lr, float b1, float b2,
              float eps, float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(blocks_of(n, 256)), dim3(256), 0, s, w, m1, m2, g, static_cast<u16 *>(wb), n,
                       lr, b1, b2, eps, c1, c2);
}

static dim3 interior_grid(const Geo &g) { return dim3((g.nz + 63) / 64, (g.ny + 3) / 4, g.nx); }

void net_store(const Geo &g, C5 U, C5 Uq, const double *J, float *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_store, interior_grid(g), dim3(64, 4), 0, s, g, U, Uq, J, S);
}

void net_gather(const float *S, const float *rs, int cap, long vol, int nx, int ny, int nz, const int *org,
                int whole_rec, const NetDims &m, int halo, void *X, float *T, unsigned char *mask, hipStream_t s)
{
    const long nv = (long)m.B * m.D * m.H * m.W;
    hipLaunchKernelGGL(k_gather, dim3(blocks_of(nv, 256)), dim3(256), 0, s, S, rs, cap, vol, nx, ny, nz, org,
                       whole_rec, m, halo, static_cast<u16 *>(X), T, mask);
}

void net_pack_input(const float *x, void *X, long nv, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_of(nv, 256)), dim3(256), 0, s, x, static_cast<u16 *>(X), nv);
}

// rs[NET_NSCALE] = the inverse scales of record a given the two before it
// (b, c; null where they do not exist)
void net_record_scales(const float *Sa, const float *Sb, const float *Sc, long vol, double *partial, int nblk,
                       float *rs, hipStream_t s)
{
    const int nprev = Sb ? (Sc ? 2 : 1) : 0;
    hipLaunchKernelGGL(k_record_sums, dim3(nblk), dim3(256), 0, s, Sa, Sb ? Sb : Sa, Sc ? Sc : (Sb ? Sb : Sa), vol,
                       partial);
    hipLaunchKernelGGL(k_record_scales, dim3(1), dim3(64), 0, s, partial, nblk, vol, nprev, rs);
}

void net_guess(const Geo &g, C5 U, C5 Uq, const double *J, const float *prev, const float *Y, const float *rs, F5 U_,
               hipStream_t s)
{
    hipLaunchKernelGGL(k_guess, interior_grid(g), dim3(64, 4), 0, s, g, U, Uq, J, prev, Y, rs, U_);
}

}  // namespace cgli
