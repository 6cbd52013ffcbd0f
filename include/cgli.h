/*
 * cgli.h -- C ABI of the MI355X (gfx950) residual + LU-SGS hot path.
 *
 * Drop-in for the reference solver's hot-path routines (main.h of
 * ricklee209/CGLI_Fu_Flow).  Fields live in HBM inside a context and keep
 * the reference's array layout: a field is X_np x Y_m x Z_m doubles,
 * index (i*Y_m + j)*Z_m + k with k fastest (Array.h:7), X_np = nx+6,
 * Y_m = ny+4, Z_m = nz+4.  Every entry point returns 0 on success or a
 * negative CGLI_E_* code; no C++ types cross the boundary.
 *
 * One context = one slab on one device, one HIP stream; entry points of
 * one context must not be called concurrently from several threads.
 */
#ifndef CGLI_H
#define CGLI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGLI_OK 0
#define CGLI_E_ARG (-1)        /* bad argument or unknown field */
#define CGLI_E_HIP (-2)        /* a HIP runtime call failed */
#define CGLI_E_NOMEM (-3)      /* device allocation failed */
#define CGLI_E_NONFINITE (-4)  /* the residual norm is NaN or Inf */
#define CGLI_E_STATE (-5)      /* a required field was never set */
#define CGLI_E_NODEV (-6)      /* no HIP device */
#define CGLI_E_TIMEOUT (-7)    /* a bounded device wait gave up */
#define CGLI_E_IO (-8)         /* a restart file could not be opened, read or written */
#define CGLI_E_SIZE (-9)       /* the grid exceeds a fixed maximum (cgli_dplusgs_noalloc) */

typedef struct cgli_ctx cgli_ctx;
typedef struct cgli_net cgli_net;
typedef struct cgli_graph cgli_graph;

/* Runtime form of the reference's compile-time configuration
 * (Resolution.h, prm.h, Viscous_terms.h, Mpi_division.h). */
typedef struct {
    int nx, ny, nz;          /* interior cells of the slab: gcount, Y_out, Z_out */
    int first, last;         /* slab has the low / high x end (myid==0 / myid==nproc-1) */
    int gstart, gend0;       /* global index of the slab's first / last interior cell */
    int nx_global;           /* X_out of the whole problem */
    int nx_inlet, nx_outlet; /* absorbing-zone widths (Resolution.h:1-2) */
    int sponge;              /* 1: absorbing-zone sources of Dual_time_stepping.cpp:532-664 */
    int xi_mode;             /* Flux_Y/Z pressure-jump normal: 0 = as the reference (none), 1 = face normal */
    int x_periodic;          /* 1: one periodic slab: x ghosts filled periodically, zero implicit x ghosts */
    int precision;           /* 0: FP64 residual kernels, 1: FP32 residual and filter kernels (FP64 update) */
    double dxi, deta, dzeta; /* deltaXI, deltaET, deltaZT */
    double high;             /* prm.h:16 */
    double K, R;             /* ratio of specific heats, gas constant */
    double Cv, mu_L, Pr_L;   /* Viscous_terms.h:3-5 */
    double suth_T, suth_S;   /* Sutherland reference temperature and constant */
    double rho_ref;          /* buoyancy / absorbing-zone reference density */
    double grav;             /* buoyancy acceleration (0: none) */
    double f_const;          /* constant streamwise body force per unit volume */
    double f_dplusgs;        /* body force of the Jacobi variant (prm.h:11) */
    double E_ref;            /* absorbing-zone target energy */
} cgli_config;

/* Field ids.  Vector fields occupy consecutive ids (component 1 first). */
enum {
    CGLI_U_ = 0,             /* U1_..U5_: Q/J at the new time level (5) */
    CGLI_U = 5,              /* U1..U5: level n (5) */
    CGLI_UQ = 10,            /* U1q..U5q: level n-1 (5) */
    CGLI_J = 15,             /* cell Jacobian */
    CGLI_MET = 16,           /* xidx xidy xidz etdx etdy etdz ztdx ztdy ztdz (9) */
    CGLI_J_U = 25, CGLI_MET_U = 26,   /* x-face J_u and 9 metrics */
    CGLI_J_V = 35, CGLI_MET_V = 36,   /* y-face */
    CGLI_J_W = 45, CGLI_MET_W = 46,   /* z-face */
    CGLI_INFX = 55, CGLI_INFY = 60, CGLI_INFZ = 65,  /* inviscid face fluxes (5 each) */
    CGLI_VF = 70,            /* vF2..vF5 (4) */
    CGLI_MR = 74, CGLI_NL = 79, CGLI_NR = 84,        /* implicit work arrays (5 each) */
    CGLI_RES = 89,           /* Residual1..5 (5) */
    CGLI_EPX = 94, CGLI_EPY = 95, CGLI_EPZ = 96,     /* filter sensor outputs */
    CGLI_YPOINT = 97,        /* grid y coordinate (Filter input) */
    CGLI_NFIELDS = 98
};

/* Reference constants (prm.h, Viscous_terms.h, Resolution.h deltas) for a
 * single slab covering the whole problem. */
void cgli_config_reference(cgli_config *cfg, int nx, int ny, int nz);

/* Context: allocates nothing up front; fields are allocated on first use. */
int cgli_create(const cgli_config *cfg, int device, cgli_ctx **out);
int cgli_destroy(cgli_ctx *ctx);
const char *cgli_error_string(int code);

/* Whole-field copies in the reference layout (X_np*Y_m*Z_m doubles). */
int cgli_set_field(cgli_ctx *ctx, int field, const double *host);
int cgli_get_field(cgli_ctx *ctx, int field, double *host);
/* Device address of a field (allocating it), for zero-copy interop.  Under
 * LU-SGS schedule 3 the increments NL equal NR on the interior after every
 * dual-time step but are written only when asked for: by cgli_get_field or
 * by this call, so a caller that keeps NL's address across steps asks for it
 * again before reading NL on the device. */
int cgli_field_device_ptr(cgli_ctx *ctx, int field, void **dptr);
int cgli_synchronize(cgli_ctx *ctx);
/* device memory of one device in bytes (free, total), for footprint reports */
int cgli_device_memory(int device, size_t *free_bytes, size_t *total_bytes);
/* Measured HBM bandwidth (no counterpart in the reference): copies of
 * `bytes` between two device buffers, `reps` times after one warm-up;
 * ms[0] per copy with one double per lane (the kernel of the time-level
 * shift), ms[1] with 16 bytes per lane on a grid-stride loop.  Bandwidth =
 * 2 * bytes / time.  Allocates and frees the two buffers. */
int cgli_stream_copy(int device, size_t bytes, int reps, double ms[2]);

/* ---- hot path: each replaces the reference routine named ------------- */
/* Flux_X.cpp:16-360 / Flux_Y.cpp:15-415 / Flux_Z.cpp:15-360:
 * reads U_, J, J_u|J_v|J_w and the face metrics; writes inFx|inFy|inFz.
 * The ML/MR scratch of the reference is not materialised. */
int cgli_flux_x(cgli_ctx *ctx, double Ep, double e);
int cgli_flux_y(cgli_ctx *ctx, double Ep, double e);
int cgli_flux_z(cgli_ctx *ctx, double Ep, double e);
/* Roe dissipation scaling of the three fluxes (off by default).  The live
 * reference fluxes ignore their Ep argument and EpX/EpY/EpZ; the older flux
 * files scale the dissipation Fav of every face by Ep
 * (Flux_X~~.cpp:552-556).  mode 0: unscaled (the live reference); 1: the
 * Ep argument of each flux call scales it; 2: the filter sensor's face
 * value EpX/EpY/EpZ (from the last cgli_filter) scales it, which is how the
 * sensor was meant to act.  Flux_Y's wall-face pressure flux keeps its own
 * Ep term (Flux_Y.cpp:400-408) in every mode. */
int cgli_set_dissipation_scaling(cgli_ctx *ctx, int mode);
/* Viscous_terms.cpp:15-908: reads U_, J and the face metrics; writes vF. */
int cgli_viscous(cgli_ctx *ctx);
/* Filter.cpp:12-1720: the multi-scale filter sensor.  Reads U_ (U1..U4),
 * J and Y_point; writes EpX/EpY/EpZ = E_high where the h/2h filtered
 * kinetic-energy ratio is >= Roe_criterion, else E_low (the reference's
 * defaults are 0.007, 1, 0.1, main.cpp:60-62).  Needs nx >= 3. */
int cgli_filter(cgli_ctx *ctx, double Roe_criterion, double E_high, double E_low);
/* (with cfg.precision = 1 the filter passes, energies and sensor ratio run
 * in FP32; faces whose ratio lies within FP32 rounding of the criterion may
 * then take the other value) */
/* The filtered kinetic energies of the last cgli_filter call, per cell in the
 * field layout: E_h (Filter.cpp's h scale) and E_2h (the 2h scale), whose
 * face ratio the sensor compares with Roe_criterion (no counterpart in the
 * reference's interface: its Filter keeps them in local arrays).  Cells the
 * sensor does not read may hold anything. */
int cgli_filter_energies(cgli_ctx *ctx, double *Eh, double *E2h);
/* Statistic.cpp:130-164: the mean and raw second-moment profiles U V W T
 * UU VV WW TT UV UT VT of U_, averaged over x and z (the channel's
 * homogeneous directions) for every interior j; out[11*ny] on the host,
 * out[q*ny + (j-2)].  FP64 whatever the residual precision. */
int cgli_profiles(cgli_ctx *ctx, double *out);
/* Hyperplane.cpp:18-81: writes the reference's 1-based plane lists to host
 * arrays Nijk[(gend+ny+nz+8)][2] and ijk[nx*ny*nz+1][3]. */
int cgli_hyperplane(cgli_ctx *ctx, int *Nijk, int *ijk);
/* Dual_time_stepping.cpp:18-1794: one inner iteration of the implicit
 * update (RHS, Gamma^-1, diagonal, hyperplane LU-SGS, update of U_, norm).
 * er[5] receives the reference's residual norms.  `sweeps` symmetric
 * sweeps (the reference runs 2; on one slab the second repeats the first).
 * write_residual != 0 also stores the squared increments in Residual.
 * On an interior slab of a multi-slab problem (first = last = x_periodic =
 * 0) the x ghost planes of NL and NR are used as the caller left them: the
 * reference exchanges them with the neighbouring ranks before each sweep
 * (Dual_time_stepping.cpp:942-976), which this single-device library does
 * not do, so with sweeps > 1 such a slab differs from a multi-rank run. */
int cgli_dual_time_stepping(cgli_ctx *ctx, double deltaT, double e, double er[5], int sweeps,
                            int write_residual);
/* The same in two halves, so that an inner iteration can be queued (or
 * captured into a graph) without waiting for the device: _async queues
 * every launch of cgli_dual_time_stepping up to the copy of the norms to
 * the host; _wait waits for them and returns er[5] and the error checks
 * (CGLI_E_NONFINITE, CGLI_E_TIMEOUT); CGLI_E_STATE when nothing is queued. */
int cgli_dual_time_stepping_async(cgli_ctx *ctx, double deltaT, double e, int sweeps, int write_residual);
int cgli_dual_time_stepping_wait(cgli_ctx *ctx, double er[5]);

/* ---- HIP graphs (no counterpart in the reference) ---------------------
 * cgli_graph_begin starts capturing the context's stream; the entry points
 * called until cgli_graph_end are recorded into a graph instead of running
 * (fluxes, viscous terms, filter, ghost fills, cgli_dual_time_stepping_async
 * and the channel's ghost fill).  Entry points that wait for the device
 * (field copies, cgli_dual_time_stepping, cgli_dplusgs, cgli_profiles,
 * restart files, mode setters) return CGLI_E_STATE while capturing, as does
 * any call that would allocate device memory: run the sequence once without
 * capture first.  Kernel timing does not see launches made by a graph.
 * cgli_graph_launch queues the graph on the context's stream (a deferred
 * norm in it is then collected by cgli_dual_time_stepping_wait); it returns
 * CGLI_E_STATE when the context changed since the capture (LU-SGS schedule,
 * spin limit, dissipation scaling, a larger reduction buffer), and the
 * sequence must be captured again.  Destroy graphs before their context. */
int cgli_graph_begin(cgli_ctx *ctx);
int cgli_graph_end(cgli_ctx *ctx, cgli_graph **out);
int cgli_graph_launch(cgli_graph *graph);
int cgli_graph_destroy(cgli_graph *graph);

/* LU-SGS schedule: 0 = one launch per i+j+k hyperplane (the reference's
 * order), 1 = tiled wavefront in one launch per sweep on precomputed
 * coefficient records, 3 = the wavefront with its coefficient records,
 * right-hand sides and increments stored in the order the march reads them
 * (one contiguous block per tile and step), 4 = the wavefront on the
 * march-ordered state, each cell computing its own split coefficients and
 * diagonal and passing on the products of its increment with them (no
 * coefficient records; the default).  Schedules 1 and 3 agree with each
 * other bit for bit, and all with the plane order to rounding.  The
 * wavefronts' bounded waits give up after max_spin polls each (then
 * CGLI_E_TIMEOUT).  Schedule 2 (retired) is CGLI_E_ARG. */
int cgli_set_lusgs_mode(cgli_ctx *ctx, int mode);
/* Schedule 4's workgroup per tile: 1 = one wave marching and computing each
 * next cell's coefficients, 2-4 = one chain wave marching and 1-3 helper
 * waves computing the coefficients of the steps ahead into LDS slots (the
 * same increments bit for bit), 0 = by tile count (the default: 3 for at
 * most two tiles of 8x8 columns per compute unit, else 1).  No counterpart
 * in the reference (its sweep is one loop per hyperplane). */
int cgli_set_sweep_waves(cgli_ctx *ctx, int waves);
int cgli_set_spin_limit(cgli_ctx *ctx, int max_spin);
/* DPLUSGS.cpp:18-1238: Jacobi (DP-LUR) variant with `sweeps` sweeps (the
 * reference runs 4). */
int cgli_dplusgs(cgli_ctx *ctx, double deltaT, double er[5], double e, int sweeps);
/* No counterpart in the reference: cgli_dplusgs for callers that cannot
 * allocate device memory while they run (harnesses that forbid hipMalloc
 * after set-up, serving loops that must not touch the allocator).  Same
 * kernels and results as cgli_dplusgs; it allocates nothing: every field it
 * reads or writes (U_, U, Uq, J, the 9 cell metrics, inFx/inFy/inFz, vF,
 * MR, NL, NR, Residual) must already be on the device (cgli_set_field, or an
 * earlier call that wrote it), else CGLI_E_STATE; the norm's partial sums go
 * to a fixed 4 MiB per-device buffer inside the library, which limits the
 * slab to nx*ny*nz <= 26,843,136 cells (else CGLI_E_SIZE).  Calls on one
 * device share that buffer: calls that overlap on a device race. */
int cgli_dplusgs_noalloc(cgli_ctx *ctx, double deltaT, double er[5], double e, int sweeps);

/* ---- ghost cells ----------------------------------------------------- */
/* Y_boundary_condition.cpp:98-182: zero-gradient copy, then isothermal
 * no-slip mirror for global i in (heated_lo, heated_hi). */
int cgli_bc_y(cgli_ctx *ctx, double T_low, double T_high, int heated_lo, int heated_hi);
/* Z_boundary_condition.cpp:48-66: one periodic ghost layer in z. */
int cgli_bc_z(cgli_ctx *ctx);
/* three periodic ghost planes at each x end of a single periodic slab */
int cgli_bc_x_periodic(cgli_ctx *ctx);

/* ---- restart file ---------------------------------------------------- */
/* Output.cpp:167-284: writes the reference's restart file (LES%05d.bin;
 * the caller names it): ten blocks of nx_global*ny*nz doubles, Q = U*J of
 * U1..U5 (level n) then of U1q..U5q (level n-1), interior cells in (i, j, k)
 * order with k fastest, native byte order.  This slab writes its cells of
 * every block gstart planes in; the file is created if missing and never
 * truncated, so the slabs of one problem can write it one after another. */
int cgli_write_restart(cgli_ctx *ctx, const char *path);
/* Initial_condition.cpp:112-219 (switch_initial = 1): reads such a file into
 * U = Q/J, Uq = Q_old/J and U_ = U on the interior cells (ghosts are left to
 * the boundary fills).  Needs J.  A file shorter than the blocks is
 * CGLI_E_IO.  Neither the running maxima of stopping rule 1 nor a network's
 * ring are in the file (the reference has neither). */
int cgli_read_restart(cgli_ctx *ctx, const char *path);

/* ---- compressible turbulent channel (the benchmark workload) ---------- */
/* Periodic x and z, isothermal no-slip walls at y = +-1 with tanh
 * stretching, constant driving pressure gradient, nondimensional with
 * rho = T = U_centreline = 1.  Replaces the reference driver's setup
 * (Grid.cpp, Jacobian*.cpp, Initial_condition.cpp), which hard-codes a
 * vertical heated channel. */
typedef struct {
    double Lx, Lz;           /* domain lengths */
    double beta;             /* tanh stretching of the wall-normal grid */
    double skew;             /* a of x' = x + a sin(2 pi y) sin(2 pi z / Lz) */
    double mach;             /* centreline Mach number */
    double re_tau, re_c;     /* friction and centreline Reynolds numbers */
    double Pr, suth_S;       /* Prandtl number; Sutherland constant / T_ref */
    double cfl;              /* acoustic CFL number that sets deltaT (default 0.5):
                                deltaT = cfl / max over cells of sum_d (|U_d| + c |grad d|) / delta_d */
    double noise;            /* uniform velocity perturbation, fraction of U_c */
    double precond_e;        /* floor e of the low-Mach preconditioning parameter
                                beta = max(|u|^2 / c^2, e) passed to the flux and
                                dual-time routines (the reference's main.cpp:57 uses 1e-4
                                for its near-incompressible air); <= 0 selects mach^2 */
    unsigned long long seed; /* splitmix64 seed of the perturbation */
    int xi_mode;             /* as cgli_config.xi_mode */
} cgli_channel;

void cgli_channel_defaults(cgli_channel *ch, double re_tau);
/* configuration of a single periodic slab of nx*ny*nz cells */
int cgli_channel_config(const cgli_channel *ch, int nx, int ny, int nz, cgli_config *cfg);
/* metrics, grid and the laminar + noise initial state on all time levels;
 * *deltaT receives the CFL-limited physical step */
int cgli_channel_init(cgli_ctx *ctx, const cgli_channel *ch, double *deltaT);
/* wall, z and x ghost fills of U_ */
int cgli_channel_ghosts(cgli_ctx *ctx);
/* one physical step: the initial guess (cgli_channel_set_guess), inner
 * iterations (ghosts, fluxes, viscous terms, dual-time LU-SGS) until the
 * stopping rule (cgli_channel_set_rule; default max_v er_v / er_v(first)
 * <= tol) holds or max_inner, then the time-level shift and, with a network
 * attached, the record of the step's increment.  *inner receives the
 * iteration count. */
int cgli_channel_step(cgli_ctx *ctx, double deltaT, double tol, int max_inner, int sweeps, int *inner,
                      double er_first[5], double er_last[5]);

/* With cfg.precision = 1 the step runs its inner iterations with FP32
 * residual kernels until the drop reaches 100 x tol or stops falling, then
 * in FP64 down to tol; *inner_fp32 receives how many of the last step's
 * iterations used FP32 residuals. */
int cgli_channel_last_step(cgli_ctx *ctx, int *inner_fp32);

/* on != 0: cgli_channel_step replays each inner iteration after a step's
 * first from a HIP graph captured the second time its (residual precision,
 * deltaT, sweeps) occurs, one launch instead of about twenty; results are
 * the same bits.  Not while kernel timing is enabled (then every iteration
 * runs eagerly).  Off by default. */
int cgli_channel_set_graph(cgli_ctx *ctx, int on);

/* ---- inner-iteration rule and initial guess of the channel step ------ */
/* No counterpart in the reference's file set; the reference stops on
 * max_v er_v / E_v < 1e-4 with E_v the running maximum of er_v over the
 * whole run (main.cpp:287-304) and starts each step from U_ = U^n
 * (main.cpp:378-382).
 * rule 0: max_v er_v / er_v(first iteration of the step) <= tol (default);
 * rule 1: max_v er_v / E_v <= tol, E_v = running maximum of er_v over every
 *         inner iteration since the rule was set. */
int cgli_channel_set_rule(cgli_ctx *ctx, int rule);
/* read (set = 0) or overwrite (set = 1) the running maxima E[5] */
int cgli_channel_rule_state(cgli_ctx *ctx, double E[5], int set);
/* guess 0: U_ = U^n; 1: U_ = U^n + dq^n (linear extrapolation), with
 * dq^n = U^n - U^{n-1}; 3: U_ = U^n + 2 dq^n - dq^{n-1} (quadratic
 * extrapolation, dq^{n-1} from the attached network's ring); 2: the
 * quadratic extrapolation plus s2 y / J, y the attached network's output on
 * its latest records and s2 their scale (cgli_net_record_scales), unless the
 * residual at that guess is larger (any variable) than at guess 1, which is
 * then used.  That check costs one residual evaluation; it is made on the
 * network's first step, on every 8th step and on every step after a
 * rejection.  The prediction itself starts at the end of the previous step
 * on the network's stream, so it overlaps what the host does between
 * steps (the filter).  Modes 2 and 3 fall back to guess 1 while the network is not
 * ready (trained, three records of this grid) or the ring lacks dq^{n-1}. */
int cgli_channel_set_guess(cgli_ctx *ctx, int guess);
/* record every step's increment per unit volume, J (U^{n+1} - U^n), into
 * the network's ring (net may be NULL to detach) */
int cgli_channel_attach_net(cgli_ctx *ctx, cgli_net *net);
/* last step: *used_net = 1 if the network's guess was taken; residual sums
 * of squares per variable at the network's and the extrapolated guess, as
 * of the last step that checked it */
int cgli_channel_guess_info(cgli_ctx *ctx, int *used_net, double r_net[5], double r_ext[5]);

/* ---- learned initial guess: 3-D CNN (no counterpart in the reference) --
 * 15 -> 32 -> 32 -> 32 -> 5 channels, 3x3x3 kernels, GELU after the first
 * three, BF16 weights and activations, FP32 accumulation and pre-
 * activations, FP32 master weights trained with Adam.  Parameters, FP32, in
 * the padded layout: per layer W[Cout_pad][27][Cin_pad] then b[Cout_pad],
 * (Cin, Cout) = (15, 32) (32, 32) (32, 32) (32, 5) padded to (16, 32) (32,
 * 32) (32, 32) (32, 16); tap = (dd+1)*9 + (dh+1)*3 + (dw+1) for offsets dd,
 * dh, dw along D, H, W.  Padded entries are zero and stay zero. */
#define CGLI_NET_NPARAM 83056
int cgli_net_create(int device, unsigned long long seed, cgli_net **out);
int cgli_net_destroy(cgli_net *net);
/* FP32 master parameters: copy out (set = 0) or in (set = 1) */
int cgli_net_params(cgli_net *net, float *params, int set);
/* One pass over host arrays, channel-last: x[B][D][H][W][15] (rounded to
 * BF16 on the device), W a multiple of 16; periodic != 0 wraps D and W
 * (else zero padding).  y[..][5] (may be NULL) receives the output.  With
 * t[..][5] and mask[B*D*H*W] (1 = counted), *loss = mean over counted
 * voxels and the 5 channels of (y - t)^2 and grad[CGLI_NET_NPARAM] (may be
 * NULL) its gradient; the gradient is also kept for cgli_net_adam. */
int cgli_net_eval(cgli_net *net, int B, int D, int H, int W, int periodic, const float *x, const float *t,
                  const unsigned char *mask, float *y, double *loss, float *grad);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) with the kept gradient */
int cgli_net_adam(cgli_net *net, float lr);
/* snapshot ring of `capacity` records of an nx*ny*nz grid (FP32 dq per
 * interior cell, [cell][5], cell = (i*ny + j)*nz + k); clears it */
int cgli_net_ring(cgli_net *net, int capacity, int nx, int ny, int nz);
/* append a record from the host */
int cgli_net_push(cgli_net *net, const float *dq);
/* forget the records appended after the first nrec (at most capacity - 3
 * of them), to replay steps from an earlier state */
int cgli_net_rewind(cgli_net *net, long nrec);
/* records appended since cgli_net_ring */
int cgli_net_records(cgli_net *net, long *n);
/* inverse scales of a record held in the ring, as the kernels use them:
 * inv[0..4] = 1 / field RMS of dq_t, inv[5..9] = 1 / field RMS of d1_t =
 * dq_t - dq_{t-1}, inv[10..14] = 1 / field RMS of d2_t = d1_t - d1_{t-1}
 * (1 where an RMS is 0 or the earlier records are missing); computed on
 * the device when the record is appended */
int cgli_net_record_scales(cgli_net *net, long rec, float inv[15]);
/* Train with Adam on `steps` batches of `batch` random crops (D, H, W) =
 * (crop_d, crop_h, crop_w) of the ring.  For record t with scales s1, sd,
 * s2: input x0 = dq_t / s1, x1 = d1_t / sd, x2 = d2_t / s2, target
 * d2_{t+1} / s2, so that a zero output is the quadratic extrapolation
 * dq_{t+1} = dq_t + d1_t; crops wrap in x and z; the loss counts voxels
 * whose 4-voxel neighbourhood lies inside the crop or beyond a wall.
 * loss[steps] (may be NULL) receives each step's loss. */
int cgli_net_train(cgli_net *net, int steps, int batch, int crop_d, int crop_h, int crop_w, float lr,
                   unsigned long long seed, float *loss);
/* whole-grid pass on the latest three records (periodic in x and z):
 * x[cells][15] and y[cells][5] (either may be NULL) as the network sees
 * and returns them (x after BF16 rounding) */
int cgli_net_predict(cgli_net *net, float *x, float *y);
/* device time (ms) and count: whole-grid predictions, training steps */
int cgli_net_timing(cgli_net *net, double *predict_ms, long *n_predict, double *train_ms, long *n_train);

/* ---- timing ---------------------------------------------------------- */
/* Average device time (ms) of the last `n` launches of a kernel class
 * recorded since cgli_timing_reset; classes: 0 flux_x 1 flux_y 2 flux_z
 * 3 viscous 4 rhs_diag 5 lusgs_fwd 6 lusgs_bwd 7 update_norm 8 split_coef
 * 9 filter 10 march_scatter (LU-SGS schedules 3 and 4: increments back to
 * their fields) 11 march_gather (schedule 4: U_ and MR into march order). */
int cgli_timing_enable(cgli_ctx *ctx, int on);
/* classes timed while timing is on: bit c of mask for class c (default all).
 * Each timed launch costs two event records on the stream, so a run that
 * needs one class's figure times only that class. */
int cgli_timing_select(cgli_ctx *ctx, unsigned mask);
int cgli_timing_reset(cgli_ctx *ctx);
int cgli_timing_get(cgli_ctx *ctx, int cls, double *total_ms, long *count);

#ifdef __cplusplus
}
#endif
#endif
