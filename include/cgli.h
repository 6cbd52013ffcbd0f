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

typedef struct cgli_ctx cgli_ctx;

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
    int precision;           /* 0: FP64 residual kernels, 1: FP32 residual kernels (FP64 update) */
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
/* Device address of a field (allocating it), for zero-copy interop. */
int cgli_field_device_ptr(cgli_ctx *ctx, int field, void **dptr);
int cgli_synchronize(cgli_ctx *ctx);

/* ---- hot path: each replaces the reference routine named ------------- */
/* Flux_X.cpp:16-360 / Flux_Y.cpp:15-415 / Flux_Z.cpp:15-360:
 * reads U_, J, J_u|J_v|J_w and the face metrics; writes inFx|inFy|inFz.
 * The ML/MR scratch of the reference is not materialised. */
int cgli_flux_x(cgli_ctx *ctx, double Ep, double e);
int cgli_flux_y(cgli_ctx *ctx, double Ep, double e);
int cgli_flux_z(cgli_ctx *ctx, double Ep, double e);
/* Viscous_terms.cpp:15-908: reads U_, J and the face metrics; writes vF. */
int cgli_viscous(cgli_ctx *ctx);
/* Filter.cpp:12-1720: the multi-scale filter sensor.  Reads U_ (U1..U4),
 * J and Y_point; writes EpX/EpY/EpZ = E_high where the h/2h filtered
 * kinetic-energy ratio is >= Roe_criterion, else E_low (the reference's
 * defaults are 0.007, 1, 0.1, main.cpp:60-62).  Needs nx >= 3. */
int cgli_filter(cgli_ctx *ctx, double Roe_criterion, double E_high, double E_low);
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
 * write_residual != 0 also stores the squared increments in Residual. */
int cgli_dual_time_stepping(cgli_ctx *ctx, double deltaT, double e, double er[5], int sweeps,
                            int write_residual);
/* LU-SGS schedule: 0 = one launch per i+j+k hyperplane (the reference's
 * order), 1 = tiled wavefront in one launch per sweep (default).  Both
 * give identical bits.  The wavefront's bounded waits give up after
 * max_spin polls each (then CGLI_E_TIMEOUT). */
int cgli_set_lusgs_mode(cgli_ctx *ctx, int mode);
int cgli_set_spin_limit(cgli_ctx *ctx, int max_spin);
/* DPLUSGS.cpp:18-1238: Jacobi (DP-LUR) variant with `sweeps` sweeps (the
 * reference runs 4). */
int cgli_dplusgs(cgli_ctx *ctx, double deltaT, double er[5], double e, int sweeps);

/* ---- ghost cells ----------------------------------------------------- */
/* Y_boundary_condition.cpp:98-182: zero-gradient copy, then isothermal
 * no-slip mirror for global i in (heated_lo, heated_hi). */
int cgli_bc_y(cgli_ctx *ctx, double T_low, double T_high, int heated_lo, int heated_hi);
/* Z_boundary_condition.cpp:48-66: one periodic ghost layer in z. */
int cgli_bc_z(cgli_ctx *ctx);
/* three periodic ghost planes at each x end of a single periodic slab */
int cgli_bc_x_periodic(cgli_ctx *ctx);

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
    double cfl;              /* physical CFL number that sets deltaT */
    double noise;            /* uniform velocity perturbation, fraction of U_c */
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
/* one physical step: inner iterations (ghosts, fluxes, viscous terms,
 * dual-time LU-SGS) until max_v er_v / er_v(first) <= tol or max_inner,
 * then the time-level shift.  *inner receives the iteration count. */
int cgli_channel_step(cgli_ctx *ctx, double deltaT, double tol, int max_inner, int sweeps, int *inner,
                      double er_first[5], double er_last[5]);

/* With cfg.precision = 1 the step runs its inner iterations with FP32
 * residual kernels until the drop reaches 100 x tol or stops falling, then
 * in FP64 down to tol; *inner_fp32 receives how many of the last step's
 * iterations used FP32 residuals. */
int cgli_channel_last_step(cgli_ctx *ctx, int *inner_fp32);

/* ---- timing ---------------------------------------------------------- */
/* Average device time (ms) of the last `n` launches of a kernel class
 * recorded since cgli_timing_reset; classes: 0 flux_x 1 flux_y 2 flux_z
 * 3 viscous 4 rhs_diag 5 lusgs_fwd 6 lusgs_bwd 7 update_norm 8 split_coef
 * 9 filter. */
int cgli_timing_enable(cgli_ctx *ctx, int on);
int cgli_timing_reset(cgli_ctx *ctx);
int cgli_timing_get(cgli_ctx *ctx, int cls, double *total_ms, long *count);

#ifdef __cplusplus
}
#endif
#endif
