// cgli_internal.hpp -- shared between the kernels and the C-ABI layer.
//
// Field layout is the reference's (Array.h:7): (i*Ym + j)*Zm + k, k
// fastest.  Thread x runs along k so a wavefront reads consecutive doubles.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "cgli.h"

namespace cgli {

// Slab geometry in the reference's index names (Resolution.h:15-29):
// interior i=3..gend, j=2..NY, k=2..NZ; wall/periodic ghosts j=1,NYY and
// k=1,NZZ; x ghosts i=0..2 and gend+1..gend+3.
struct Geo {
    long Ym, Zm, N;
    int nx, ny, nz;
    int gend, NY, NZ, NYY, NZZ;
    int first, last;
    int xper;   // single periodic slab
};

struct Phys {
    double K, R, Cv, mu_L, Pr_L, suth_T, suth_S;
    double dxi, deta, dzeta;
    double rho_ref, grav, f_const, E_ref, high;
    int fp32;   // residual kernels (fluxes, viscous terms) in FP32 arithmetic
};

struct F5 { double *v[5]; };
struct C5 { const double *v[5]; };
struct C4 { const double *v[4]; };
struct F4 { double *v[4]; };
struct M9 { const double *m[9]; };

// x / d from r = 1 / d (the correctly rounded reciprocal): the product q and
// one fused multiply-add correction, q + (x - q d) r, is the correctly
// rounded quotient -- the division's own bits -- barring underflow and
// overflow (P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Develop. 34 (1990); J.-M. Muller
// et al., Handbook of Floating-Point Arithmetic, on division by
// Newton-Raphson iteration), in any binary format with a fused multiply-add.
// Three instructions instead of a division's ten when several quotients
// share a divisor.
__host__ __device__ __forceinline__ double div_r(double x, double d, double r)
{
    const double q = x * r;
    return fma(fma(-q, d, x), r, q);
}
// the same in binary32
__host__ __device__ __forceinline__ float div_r(float x, float d, float r)
{
    const float q = x * r;
    return fmaf(fmaf(-q, d, x), r, q);
}

__host__ __device__ inline long ix(const Geo &g, int i, int j, int k)
{
    return ((long)i * g.Ym + j) * g.Zm + k;
}

// launchers (kernels.hip)
// epS scales every face's Roe dissipation, or EpF[face] does when EpF is
// not null (1 and null: the live reference flux)
void launch_flux_x(const Geo &g, const Phys &p, double e, C5 U, const double *J, const double *Jf, M9 mf,
                   F5 F, double epS, const double *EpF, hipStream_t s);
void launch_flux_y(const Geo &g, const Phys &p, double Ep, double e, int xi_mode, C5 U, const double *J,
                   const double *Jf, M9 mf, F5 F, double epS, const double *EpF, hipStream_t s);
void launch_flux_z(const Geo &g, const Phys &p, double e, int xi_mode, C5 U, const double *J, const double *Jf,
                   M9 mf, F5 F, double epS, const double *EpF, hipStream_t s);
void launch_viscous(const Geo &g, const Phys &p, C5 U, const double *J, const double *Ju, M9 mu,
                    const double *Jv, M9 mv, const double *Jw, M9 mw, F4 vF, hipStream_t s);
// zeroes counters[0..1] too when counters is not null
void launch_zero_increment_ghosts(const Geo &g, F5 NL, F5 NR, int *counters, hipStream_t s);
void launch_sponge(const Geo &g, const Phys &p, double e, int gstart, int gend0, int nx_global, int nx_inlet,
                   int nx_outlet, C5 U, const double *J, const double *etdy, F5 Res, double *scratch,
                   hipStream_t s);
// write_inc = 0: MR only, not the starting increments NL = NR = D MR (the
// march-ordered sweep keeps its own copies and its scatter writes both)
void launch_rhs_diag(const Geo &g, const Phys &p, double deltaT, double e, int with_res, C5 U_, C5 U, C5 Uq,
                     const double *J, M9 met, C5 Fx, C5 Fy, C5 Fz, C4 vF, C5 Res, F5 MR, F5 NL, F5 NR,
                     hipStream_t s, int write_inc = 1);
void launch_lusgs_plane(const Geo &g, const Phys &p, double deltaT, double e, int forward, int m,
                        int first_cell, int count, const int *ijk, C5 U_, const double *J, M9 met, C5 MR,
                        F5 NL, F5 NR, hipStream_t s);
void launch_dplur_sweep(const Geo &g, const Phys &p, double deltaT, double e, C5 U_, const double *J, M9 met,
                        C5 MR, C5 NLold, F5 NRnew, hipStream_t s);
void launch_copy_interior(const Geo &g, C5 src, F5 dst, hipStream_t s);
// returns the number of partial blocks written to `partials` (5 per block)
int launch_update_norm(const Geo &g, const Phys &p, F5 U_, const double *J, C5 NR, F5 Res, int write_res,
                       int dplusgs, double *partials, hipStream_t s);
// er_dev may be a device address of pinned host memory; with counters,
// counters[0..1] are copied to flags
void launch_finish_norm(int nblocks, const double *partials, double DN, int squared_twice, double *er_dev,
                        hipStream_t s, const int *counters = nullptr, int *flags = nullptr);
void launch_bc_y(const Geo &g, const Phys &p, int gstart, F5 U, const double *J, double T_low,
                 double T_high, int heated_lo, int heated_hi, hipStream_t s);
void launch_bc_z(const Geo &g, F5 U, hipStream_t s);
void launch_bc_x_periodic(const Geo &g, F5 U, hipStream_t s);
void launch_pack_q(const Geo &g, const double *U, const double *J, double *out, hipStream_t s);
void launch_unpack_q(const Geo &g, const double *in, const double *J, double *U, double *U2, hipStream_t s);
// zero fill by a kernel (safe inside a captured graph, unlike a memset node)
void launch_zero(double *p, long n, hipStream_t s);
// copy plane 0 of dst to planes 1..nplanes-1
void launch_replicate_planes(double *dst, long plane, int nplanes, hipStream_t s);
void launch_copy_field(double *dst, const double *src, long n, hipStream_t s);
// n even; blocks of 256 lanes, grid-stride over n/2 16-byte elements
void launch_stream_copy(double *dst, const double *src, long n, int blocks, hipStream_t s);

// filter sensor (filter.hip); work holds filter_scratch_doubles(g) doubles
long filter_scratch_doubles(const Geo &g);
// fp32: the passes, energies and sensor ratio in FP32 (work then holds floats)
void launch_filter(const Geo &g, double crit, double E_high, double E_low, C5 U_, const double *J,
                   const double *Y, double *work, int fp32, double *EpX, double *EpY, double *EpZ, hipStream_t s);

// wall-normal profiles (stats.hip): out = 11 x ny doubles on the device
void launch_profiles(const Geo &g, const Phys &p, C5 U, const double *J, double *out, hipStream_t s);

// wavefront LU-SGS (implicit.hip)
constexpr int WAVE_TJ = 4, WAVE_TK = 16;
// tile of the wavefront on march-ordered records (one wave: TJ*TK = 64)
#ifndef CGLI_MARCH_TJ
#define CGLI_MARCH_TJ 8
#define CGLI_MARCH_TK 8
#endif
constexpr int MARCH_TJ = CGLI_MARCH_TJ, MARCH_TK = CGLI_MARCH_TK;
// march steps each tile starts later than its dependencies require, so that
// its exchange polls find the upstream entries already written (measured at
// 128^3: 0 -> 722/694 us per forward/backward sweep, 1 -> 686/651,
// 2 -> 645/609, 3 -> 644/612, 5 -> 712/681)
#ifndef CGLI_MARCH_SLACK
#define CGLI_MARCH_SLACK 2
#endif
constexpr int MARCH_SLACK = CGLI_MARCH_SLACK;
int wave_tiles(const Geo &g, int tj, int tk, std::vector<int> &order);
void launch_split_coef(const Geo &g, const Phys &p, double deltaT, double e, C5 U_, const double *J, M9 met,
                       void *coef, double *dgn, int fp32, hipStream_t s);
// doubles of the coefficient blocks of the wavefront sweep
long coef_doubles(const Geo &g);
long diag_doubles(const Geo &g);
// coef: coefficient blocks, xbuf: 5 fields (tile-edge exchange),
// trash: 256 doubles per tile
// before each sweep: exchange sentinels and the tile ticket
void launch_lusgs_wave_prep(const Geo &g, double *xbuf, int *counters, hipStream_t s);
// before a march sweep: sentinels into the exchange entries it polls, ticket
// counters[0] to zero; xcol: the buffer of the k-edge entries when it is not
// xbuf (schedule 4 exchanges different j and k products)
void launch_march_prep(const Geo &g, int forward, double *xbuf, int *counters, hipStream_t s,
                       double *xcol = nullptr);
// coef: split blocks (float when fp32), dgn: FP64 diagonal blocks
void launch_lusgs_wave(const Geo &g, int forward, const void *coef, const double *dgn, int fp32, C5 R, F5 X,
                       double *xbuf, double *trash, const int *order, int ntiles, int *counters, int max_spin,
                       hipStream_t s);
void launch_reset_ints(int *p, int n, hipStream_t s);
// march layout (implicit.hip): coefficient records, right-hand sides and
// increments in the order the wavefront reads them; sizes in CT pairs, FP64
// pairs and doubles
long march_coef_pairs(const Geo &g);
long march_diag_pairs(const Geo &g);
long march_field_doubles(const Geo &g);
bool march_fits(const Geo &g, int fp32);
void launch_split_march(const Geo &g, const Phys &p, double deltaT, double e, C5 U_, const double *J, M9 met,
                        C5 MR, void *coefM, double *dgnM, double *mrM, int fp32, hipStream_t s);
void launch_lusgs_march(const Geo &g, int forward, const void *coefM, const double *dgnM, int fp32,
                        const double *Rm, double *Xm, C5 Xg, double *xbuf, double *trash, const int *order,
                        int ntiles, int *counters, int max_spin, int slack, hipStream_t s);
// NR (march order) into the fields NR and NL: the end of a sweep pair
// NL with null pointers: NR only
void launch_march_scatter(const Geo &g, const double *nrM, F5 NL, F5 NR, hipStream_t s);
// schedule 4 (implicit.hip): march-ordered pairs of the geometry (J and 9
// metrics, 5 pairs per slot) and of five-value fields (3 pairs per slot)
long march4_geo_pairs(const Geo &g);
long march4_field_pairs(const Geo &g);
// U_ and R (when Rm) and the geometry (when Gm) into march order
void launch_gather4(const Geo &g, C5 U_, C5 R, const double *J, M9 met, void *Um, void *Rm, void *Gm, hipStream_t s);
// march-ordered increments Xm into NR (and NL unless its pointers are null)
void launch_scatter4(const Geo &g, const void *Xm, F5 NL, F5 NR, hipStream_t s);
// schedule 4's update and norm from march-ordered U_ (Um) and NR (Xm);
// returns the number of partial-sum blocks (update4_blocks)
long update4_blocks(const Geo &g);
int launch_update4(const Geo &g, const Phys &p, F5 U_, const double *J, void *Um, const void *Xm, F5 Res,
                   int write_res, double *partials, hipStream_t s);
// bytes of one of the two tile-edge exchange record buffers of a sweep (the
// j products, then the k products); the sweep needs 2 * march4_rec_bytes
// below 2 GiB
long march4_rec_bytes(const Geo &g);
// the sentinel into both record buffers and the sweep's tile ticket to zero
void launch_rec4_prep(void *rec, long rec_bytes, int *counters, hipStream_t s);
// one sweep: right-hand sides Rm (forward MR, backward NL) into increments
// Xm; U_, J, met and the increment field Xg supply the ghost cell upstream
// in i; rec: the two exchange record buffers (launch_rec4_prep before each
// sweep); slack < 0: the schedule's own lead per hop; waves: workgroup waves
// per tile (1, or a chain wave and 1-3 helpers)
// Gc not null: the metrics do not vary along i and Gc holds each tile
// lane's column geometry (launch_gather_cols4); Gm is then not read
void launch_lusgs_fused(const Geo &g, const Phys &p, double deltaT, double e, int forward, const void *Gm,
                        const void *Gc, const void *Um, const void *Rm, void *Xm, C5 U_, const double *J, M9 met, C5 Xg,
                        void *rec, long rec_bytes, const int *order, int ntiles, int *counters,
                        int max_spin, int slack, int waves, hipStream_t s);
// *flag = 1 where an interior cell's J or metrics differ from its column's
// at the first interior plane (the caller zeroes *flag)
void launch_geo_icheck(const Geo &g, const double *J, M9 met, int *flag, hipStream_t s);
// pairs of the per-column geometry array of the GC sweep, and its gather
long march4_col_pairs(const Geo &g);
void launch_gather_cols4(const Geo &g, const double *J, M9 met, void *Gc, hipStream_t s);


// learned initial guess (cnn.hip): see that file for layouts
struct NetDims { int B, D, H, W; };
constexpr int NET_NSCALE = 15;   // inverse scales per record: 1/s1, 1/sd, 1/s2
int net_conv(int cin_pad, int cout_pad, bool act, const void *X, const void *W, const float *bias, float *Z, void *A,
             const NetDims &m, int cout, int per, hipStream_t s);
void net_gelu_back(const float *dA, const float *Z, void *dZ, float *dZf, long n, hipStream_t s);
void net_to_bf16(const float *x, void *y, long n, hipStream_t s);
void net_transpose_w(const void *W, void *WT, int cin_pad, int cout_pad, hipStream_t s);
int net_wgrad(int cin_pad, int cout_pad, const void *dZ, const void *X, const NetDims &m, int nblk, float *part,
              int cout, int cin, float *dW, hipStream_t s);
constexpr int NET_BIAS_BLOCKS = 256;
// cpad 16 or 32; part holds NET_BIAS_BLOCKS x cpad doubles
void net_bias_grad(const float *dZ, long nvox, int cpad, int cout, float *db, double *part, hipStream_t s);
void net_mse(const float *Y, const float *T, const unsigned char *mask, long nvox, float inv, float *dY,
             double *partial, int nblk, double *loss, double loss_scale, hipStream_t s);
void net_adam(float *w, float *m1, float *m2, const float *g, void *wb, long n, float lr, float b1, float b2,
              float eps, float c1, float c2, hipStream_t s);
void net_store(const Geo &g, C5 U, C5 Uq, const double *J, float *S, hipStream_t s);
void net_gather(const float *S, const float *rs, int cap, long vol, int nx, int ny, int nz, const int *org,
                int whole_rec, const NetDims &m, int halo, void *X, float *T, unsigned char *mask, hipStream_t s);
void net_pack_input(const float *x, void *X, long nv, hipStream_t s);
void net_record_scales(const float *Sa, const float *Sb, const float *Sc, long vol, double *partial, int nblk,
                       float *rs, hipStream_t s);
void net_guess(const Geo &g, C5 U, C5 Uq, const double *J, const float *prev, const float *Y, const float *rs, F5 U_,
               hipStream_t s);
// sum over the interior cells of f_q^2 for q < 5 into out_dev[5] (stats.hip)
void launch_sumsq5(const Geo &g, C5 f, double *partials, double *out_dev, hipStream_t s);

// context access for host modules other than capi.cpp
int ctx_field(cgli_ctx *c, int id, double **dptr);   // allocates on first use
hipStream_t ctx_stream(cgli_ctx *c);
const cgli_config *ctx_config(cgli_ctx *c);
void ctx_mark(cgli_ctx *c, int id, int n);
void ctx_set_fp32(cgli_ctx *c, bool on);          // residual kernels in FP32 from now on
void ctx_set_last_fp32(cgli_ctx *c, int n);       // FP32 inner iterations of the last step
const Geo &ctx_geo(cgli_ctx *c);
// residual MR of the current U_ (rhs_diag after the caller's flux and
// viscous calls); out[5] = its interior sum of squares per variable
int ctx_rhs_sumsq(cgli_ctx *c, double deltaT, double e, double out[5]);
// one inner iteration (cgli_dual_time_stepping); rhs_ready: the last call
// on U_ was ctx_rhs_sumsq, whose right-hand side is reused
int ctx_dts(cgli_ctx *c, double deltaT, double e, double er[5], int sweeps, int write_residual, bool rhs_ready);
// the same in two halves: queue everything up to the norm's copy to the
// host, then wait for it and check it
// u4_current: U_'s interior is unchanged since the previous call's update
// (the caller's promise; an inner iteration after the first of one step)
int ctx_dts_async(cgli_ctx *c, double deltaT, double e, int sweeps, int write_residual, bool rhs_ready,
                  bool u4_current = false);
int ctx_dts_wait(cgli_ctx *c, double er[5]);
bool ctx_timing_on(cgli_ctx *c);
// per-context state of the channel driver (channel.cpp)
struct ChannelState {
    double e = 1e-4;         // preconditioning floor of the step's routines (set by cgli_channel_init)
    int rule = 0;            // 0: per-step drop, 1: running maximum
    double E[5] = {0, 0, 0, 0, 0};
    int guess = 0;           // 0: U^n, 1: linear, 2: network, 3: quadratic extrapolation
    cgli_net *net = nullptr;
    int used_net = 0;        // last step: 1 if the network's guess was taken
    double r_net[5] = {0, 0, 0, 0, 0}, r_ext[5] = {0, 0, 0, 0, 0};
    int since_check = 0;     // steps since the network's guess was last checked
    bool verify = true;      // check at the next step (first use, after a rejection)
    // inner iterations replayed from captured graphs, one per (residual
    // precision, deltaT, sweeps, first or later iteration of a step); a key
    // runs eagerly the first time
    struct GraphSlot {
        int fp32;
        double dt;
        int sweeps;
        int later;
        cgli_graph *g;
    };
    bool graph = false;
    std::vector<GraphSlot> graphs;
};
void channel_drop_graphs(ChannelState &cs);
ChannelState &ctx_channel(cgli_ctx *c);
// network side of the channel driver (learned.cpp)
int net_record(cgli_net *net, cgli_ctx *c);                 // append U - Uq to the ring
int net_ready(cgli_net *net, cgli_ctx *c);                  // trained, 3 records of this grid
// whole-grid prediction on the latest records: Y [cell][16] and the
// latest record's inverse scales rs[NET_NSCALE], both on the device
int net_predict(cgli_net *net, hipStream_t s, const float **Y, const float **rs);
// start the prediction for the next step on the network's stream, after the
// context stream's work so far (the latest record); it runs while the host
// does other work (the filter) and the next guess takes it
int net_predict_ahead(cgli_net *net, cgli_ctx *c);
// the prediction started ahead for the current records, if any: the
// context stream waits for it; false when there is none
bool net_take_ahead(cgli_net *net, cgli_ctx *c, const float **Y, const float **rs);
// the ring's record before the latest ([cell][5], device), or null when
// the ring does not hold two records of this grid
const float *net_prev_record(cgli_net *net, cgli_ctx *c);

}  // namespace cgli
