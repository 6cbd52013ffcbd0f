/*
 * cgli_oracle.h -- CPU oracle for the residual + LU-SGS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's CPU routines (ricklee209/CGLI_Fu_Flow), written from its
 * arithmetic so the GPU path can be checked against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product library never links or calls it.
 *
 * Layout: every field is a double array laid out exactly like the
 * reference's `double (*)[Y_m][Z_m]` arrays (Array.h:7): index
 * ((i*Y_m)+j)*Z_m+k, k fastest, with X_np = nx+6, Y_m = ny+4, Z_m = nz+4
 * (main.cpp:85, Resolution.h:11-13).  Interior cells are i=3..gend
 * (gend = nx+2, Mpi_division.h:27), j=2..NY (NY = ny+1), k=2..NZ.
 *
 * Rank semantics: `first`/`last` reproduce the reference's myid==0 /
 * myid==nproc-1 branches; gstart/gend0/nx_global reproduce
 * Mpi_division.h's global offsets.
 */
#ifndef CGLI_ORACLE_H
#define CGLI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;          /* interior cells of this slab: gcount, Y_out, Z_out */
    int first, last;         /* slab touches the low / high x end (myid==0 / myid==nproc-1) */
    int gstart, gend0;       /* global index of the first / last interior cell */
    int nx_global;           /* X_out of the whole problem */
    int nx_inlet, nx_outlet; /* sponge widths (Resolution.h:1-2) */
    int sponge;              /* 1: reference absorbing zones (Dual_time_stepping.cpp:532-664) */
    int xi_mode;             /* Flux_Y/Z Fav2..4 normal: 0 = zero (UB compat), 1 = face unit normal */
    int x_periodic;          /* one periodic slab: implicit x ghosts are zero as at nproc=1 */
    int precision;           /* unused by the oracle (FP64 only) */
    double dxi, deta, dzeta; /* deltaXI, deltaET, deltaZT */
    double high;             /* channel height used by the sponge integral (prm.h:16) */
    double K, R;             /* gamma, gas constant */
    double Cv, mu_L, Pr_L;   /* Viscous_terms.h:3-5 */
    double suth_T, suth_S;   /* Sutherland reference temperature and constant */
    double rho_ref;          /* buoyancy reference density (1.1842) */
    double grav;             /* buoyancy g (9.81); 0 disables */
    double f_const;          /* constant x body force added to the buoyancy term */
    double f_dplusgs;        /* DPLUSGS constant force f (prm.h:11: 0.5993) */
    double E_ref;            /* sponge target energy (253250) */
} orc_cfg;

/* Reference constants at the sizes given (everything else as in the
 * reference source). */
void orc_cfg_reference(orc_cfg *c, int nx, int ny, int nz);

/* --- deterministic inputs -------------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state);
double orc_unit(uint64_t *state);  /* uniform [0,1) from the top 53 bits */
void orc_fill(double *a, long n, uint64_t seed, double lo, double hi);
void orc_fill_state(long n, uint64_t seed, const double *J, double K, double r0, double r1,
                    double ua, double p0, double p1, double *U_[5]);

/* --- residual kernels (a1-a4) ------------------------------------------ */
void orc_flux_x(const orc_cfg *c, double Ep, double e, double *const U_[5],
                double *MR[5], double *ML[5], const double *J, const double *J_u,
                const double *const met_u[9], double *inFx[5]);
void orc_flux_y(const orc_cfg *c, double Ep, double e, double *const U_[5],
                double *MR[5], double *ML[5], const double *J, const double *J_v,
                const double *const met_v[9], double *inFy[5]);
void orc_flux_z(const orc_cfg *c, double Ep, double e, double *const U_[5],
                double *MR[5], double *ML[5], const double *J, const double *J_w,
                const double *const met_w[9], double *inFz[5]);
/* the same with the Roe dissipation of every face scaled by 1 (mode 0),
 * Ep (mode 1) or EpF[face] (mode 2) */
void orc_flux_x_scaled(const orc_cfg *c, int mode, double Ep, double e, double *const U_[5],
                       double *MR[5], double *ML[5], const double *J, const double *J_u,
                       const double *const met_u[9], double *inFx[5], const double *EpF);
void orc_flux_y_scaled(const orc_cfg *c, int mode, double Ep, double e, double *const U_[5],
                       double *MR[5], double *ML[5], const double *J, const double *J_v,
                       const double *const met_v[9], double *inFy[5], const double *EpF);
void orc_flux_z_scaled(const orc_cfg *c, int mode, double Ep, double e, double *const U_[5],
                       double *MR[5], double *ML[5], const double *J, const double *J_w,
                       const double *const met_w[9], double *inFz[5], const double *EpF);
/* vF[0..3] = vF2..vF5; scratch LL/ML/NL hold the face fluxes 2..5 */
void orc_viscous(const orc_cfg *c, double *const U_[5], double *vF[4],
                 const double *J, const double *J_u, const double *const met_u[9],
                 const double *J_v, const double *const met_v[9],
                 const double *J_w, const double *const met_w[9],
                 double *LL[5], double *ML[5], double *NL[5]);

/* --- implicit update (a6-a9, a11) -------------------------------------- */
int  orc_hyperplane_count(const orc_cfg *c);       /* number of cells listed */
void orc_hyperplane(const orc_cfg *c, int *Nijk, int *ijk);
/* Nijk has (gend+ny+nz+8) rows of 2 ints, ijk has (cells+1) rows of 3 */

void orc_dual_time_stepping(const orc_cfg *c, double deltaT, double e, double er[5],
                            const int *Nijk, const int *ijk,
                            double *U_[5], double *const U[5], double *const Uq[5],
                            const double *J, const double *const met[9],
                            double *const inFx[5], double *const inFy[5], double *const inFz[5],
                            double *const vF[4], double *MR[5], double *NR[5], double *NL[5],
                            double *Res[5], const double *etdy_v, const double *J_v,
                            int sweeps);
void orc_dplusgs(const orc_cfg *c, double deltaT, double er[5], double e,
                 double *U_[5], double *const U[5], double *const Uq[5],
                 const double *J, const double *const met[9],
                 double *const inFx[5], double *const inFy[5], double *const inFz[5],
                 double *const vF[4], double *MR[5], double *NR[5], double *NL[5],
                 double *Res[5], int sweeps);

/* --- filter sensor (a5) ------------------------------------------------ */
void orc_filter(const orc_cfg *c, double Roe_criterion, double E_high, double E_low,
                double *const U_[5], const double *Y_point,
                double *EpX, double *EpY, double *EpZ, const double *J);
/* E_h and E_2h (with their ghost cells) that the sensor compares */
void orc_filter_energy(const orc_cfg *c, double *const U_[5], const double *Y_point, const double *J,
                       double *Eh, double *E2h);

/* --- wall-normal profiles (statistics of the channel) ------------------ */
/* out[11*ny]: U V W T UU VV WW TT UV UT VT averaged over x and z per j */
void orc_profiles(const orc_cfg *c, double *const U_[5], const double *J, double *out);

/* --- metrics (a12) ----------------------------------------------------- */
/* Reference analytic tanh grid (Jacobian*.cpp), gamma1 = 2.8. met = 9
 * arrays xidx..ztdz; `cell` is the centre set used by Jacobian_v's wall
 * override. */
void orc_jacobian(const orc_cfg *c, double *J, double *met[9]);
void orc_jacobian_u(const orc_cfg *c, double *J_u, double *met_u[9]);
void orc_jacobian_v(const orc_cfg *c, double *J_v, double *met_v[9],
                    const double *J, double *const met[9]);
void orc_jacobian_w(const orc_cfg *c, double *J_w, double *met_w[9]);

/* --- ghost cells ------------------------------------------------------- */
/* Y_boundary_condition.cpp:98-182 with wall temperatures as arguments */
void orc_bc_y(const orc_cfg *c, double *U_[5], const double *J, double T_low, double T_high,
              int heated_lo, int heated_hi);
void orc_bc_z(const orc_cfg *c, double *U_[5]);           /* Z_boundary_condition.cpp:48-66 */
void orc_bc_x_periodic(const orc_cfg *c, double *U_[5]);  /* single-slab periodic halo */

#ifdef __cplusplus
}
#endif
#endif
