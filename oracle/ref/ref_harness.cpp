// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points around the reference solver's own routines, compiled from
// the reference sources where they lie (see build_ref.sh).  It lets the
// golden-vector script call Flux_X/Y/Z, Viscous_terms, Dual_time_stepping,
// Hyperplane, Jacobian*, Y_/Z_boundary_condition, Filter, and the restart
// file's writer and reader (Output, Initial_condition) on arrays it owns.  The
// reference's grid sizes are compile-time macros (Resolution.h), so one
// library is built per size; ref_dims() reports them.
#include <mpi.h>
#include "main.h"

int X_np;  // declared `extern int X_np` by every reference TU (main.h:7)

typedef double (*A3)[Y_m][Z_m];

static A3 a(double *p) { return reinterpret_cast<A3>(p); }

extern "C" {

int ref_init(void)
{
    int flag = 0;
    MPI_Initialized(&flag);
    if (!flag) MPI_Init(nullptr, nullptr);
    return 0;
}

void ref_deltas(double *out)
{
    out[0] = deltaXI; out[1] = deltaET; out[2] = deltaZT;
}

void ref_flux(int dir, int xnp, int myid, double Ep, double e, double **U_, double **MR, double **ML,
              double *J, double *Jf, double **m, double **F, double *Ep_arr)
{
    X_np = xnp;
    if (dir == 0)
        Flux_X(myid, Ep, e, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]),
               a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]),
               a(ML[0]), a(ML[1]), a(ML[2]), a(ML[3]), a(ML[4]), a(J), a(Jf),
               a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]),
               a(F[0]), a(F[1]), a(F[2]), a(F[3]), a(F[4]), a(Ep_arr));
    else if (dir == 1)
        Flux_Y(myid, Ep, e, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]),
               a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]),
               a(ML[0]), a(ML[1]), a(ML[2]), a(ML[3]), a(ML[4]), a(J), a(Jf),
               a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]),
               a(F[0]), a(F[1]), a(F[2]), a(F[3]), a(F[4]), a(Ep_arr));
    else
        Flux_Z(myid, Ep, e, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]),
               a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]),
               a(ML[0]), a(ML[1]), a(ML[2]), a(ML[3]), a(ML[4]), a(J), a(Jf),
               a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]),
               a(F[0]), a(F[1]), a(F[2]), a(F[3]), a(F[4]), a(Ep_arr));
}

// met: 9 cell metrics; mu/mv/mw: 9 face metrics each; scratch 6 x 5 arrays
void ref_viscous(int xnp, int myid, double **U_, double **vF, double *J, double **met,
                 double *J_u, double **mu, double *J_v, double **mv, double *J_w, double **mw,
                 double **LR, double **LL, double **MR, double **ML, double **NR, double **NL)
{
    X_np = xnp;
    Viscous_terms(myid, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]),
                  a(vF[0]), a(vF[1]), a(vF[2]), a(vF[3]), a(J),
                  a(met[0]), a(met[1]), a(met[2]), a(met[3]), a(met[4]), a(met[5]), a(met[6]), a(met[7]), a(met[8]),
                  a(J_u), a(mu[0]), a(mu[1]), a(mu[2]), a(mu[3]), a(mu[4]), a(mu[5]), a(mu[6]), a(mu[7]), a(mu[8]),
                  a(J_v), a(mv[0]), a(mv[1]), a(mv[2]), a(mv[3]), a(mv[4]), a(mv[5]), a(mv[6]), a(mv[7]), a(mv[8]),
                  a(J_w), a(mw[0]), a(mw[1]), a(mw[2]), a(mw[3]), a(mw[4]), a(mw[5]), a(mw[6]), a(mw[7]), a(mw[8]),
                  a(LR[0]), a(LR[1]), a(LR[2]), a(LR[3]), a(LR[4]), a(LL[0]), a(LL[1]), a(LL[2]), a(LL[3]), a(LL[4]),
                  a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]), a(ML[0]), a(ML[1]), a(ML[2]), a(ML[3]), a(ML[4]),
                  a(NR[0]), a(NR[1]), a(NR[2]), a(NR[3]), a(NR[4]), a(NL[0]), a(NL[1]), a(NL[2]), a(NL[3]), a(NL[4]));
}

void ref_hyperplane(int xnp, int myid, int *Nijk, int *ijk)
{
    X_np = xnp;
    Hyperplane(myid, reinterpret_cast<int (*)[2]>(Nijk), reinterpret_cast<int (*)[3]>(ijk));
}

void ref_dts(int xnp, int myid, double deltaT, double e, double *er, int *Nijk, int *ijk,
             double **U_, double **U, double **Uq, double *J, double **met,
             double **Fx, double **Fy, double **Fz, double **vF, double **MR, double **NR, double **NL,
             double **Res, double *etdy_v, double *J_v)
{
    X_np = xnp;
    Dual_time_stepping(myid, deltaT, e, &er[0], &er[1], &er[2], &er[3], &er[4],
                       reinterpret_cast<int (*)[2]>(Nijk), reinterpret_cast<int (*)[3]>(ijk),
                       a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]),
                       a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]),
                       a(Uq[0]), a(Uq[1]), a(Uq[2]), a(Uq[3]), a(Uq[4]), a(J),
                       a(met[0]), a(met[1]), a(met[2]), a(met[3]), a(met[4]), a(met[5]), a(met[6]), a(met[7]), a(met[8]),
                       a(Fx[0]), a(Fx[1]), a(Fx[2]), a(Fx[3]), a(Fx[4]),
                       a(Fy[0]), a(Fy[1]), a(Fy[2]), a(Fy[3]), a(Fy[4]),
                       a(Fz[0]), a(Fz[1]), a(Fz[2]), a(Fz[3]), a(Fz[4]),
                       a(vF[0]), a(vF[1]), a(vF[2]), a(vF[3]),
                       a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]),
                       a(NR[0]), a(NR[1]), a(NR[2]), a(NR[3]), a(NR[4]),
                       a(NL[0]), a(NL[1]), a(NL[2]), a(NL[3]), a(NL[4]),
                       a(Res[0]), a(Res[1]), a(Res[2]), a(Res[3]), a(Res[4]), a(etdy_v), a(J_v));
}

// which: 0 cell (J, met), 1 u faces, 2 v faces (needs the cell set in J/met), 3 w faces
void ref_jacobian(int which, int xnp, int myid, double *Jf, double **m, double *J, double **met)
{
    X_np = xnp;
    if (which == 0)
        Jacobian(myid, a(Jf), a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]));
    else if (which == 1)
        Jacobian_u(myid, a(Jf), a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]));
    else if (which == 2)
        Jacobian_v(myid, a(Jf), a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]),
                   a(J), a(met[0]), a(met[1]), a(met[2]), a(met[3]), a(met[4]), a(met[5]), a(met[6]), a(met[7]), a(met[8]));
    else
        Jacobian_w(myid, a(Jf), a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]));
}

void ref_bc_yz(int which, int xnp, int myid, double deltaT, double e, double **U_, double **U, double **Uq,
               double *etdy, double *J)
{
    X_np = xnp;
    if (which == 0)
        Y_boundary_condition(myid, deltaT, e, 104.0, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]),
                             a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]),
                             a(Uq[0]), a(Uq[1]), a(Uq[2]), a(Uq[3]), a(Uq[4]), a(etdy), a(J));
    else
        Z_boundary_condition(myid, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]), a(J));
}

void ref_filter(int xnp, int myid, double Roe_criterion, double E_high, double E_low, double **U_,
                double *Y_point, double *EpX, double *EpY, double *EpZ, double *J)
{
    X_np = xnp;
    Filter(myid, Roe_criterion, E_high, E_low, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]), a(Y_point),
           a(EpX), a(EpY), a(EpZ), a(J));
}

// Output.cpp with switch_output = 0: only the restart file LES%05d.bin,
// written into the current directory (the grid arrays are not read then).
void ref_output(int xnp, int myid, int step, double **U, double **Uq, double *J)
{
    X_np = xnp;
    Output(step, myid, 0, a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]), a(Uq[0]), a(Uq[1]), a(Uq[2]), a(Uq[3]),
           a(Uq[4]), a(J), a(J), a(J), a(J));
}

// Initial_condition.cpp with switch_initial = 1: the uniform state, then
// the interior from initial.bin in the current directory.
void ref_initial_condition(int xnp, int myid, double **U_, double **U, double **Uq, double *J)
{
    X_np = xnp;
    Initial_condition(myid, 1, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]), a(U[0]), a(U[1]), a(U[2]),
                      a(U[3]), a(U[4]), a(Uq[0]), a(Uq[1]), a(Uq[2]), a(Uq[3]), a(Uq[4]), a(J));
}

// X_out Y_out Z_out Y_m Z_m np nx_inlet nx_outlet.  MPI_prm.h is a
// function-body include in the reference (it defines locals), so it is
// included the same way here to read np.
void ref_dims(int *out)
{
#include "MPI_prm.h"
    (void)nproc;
    out[0] = X_out; out[1] = Y_out; out[2] = Z_out;
    out[3] = Y_m; out[4] = Z_m; out[5] = np;
    out[6] = nx_inlet; out[7] = nx_outlet;
}

}  // extern "C"
