// shim_probe.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Calls the reference-signature routines of include/cgli_reference_api.hpp
// the way a reference translation unit would (double (*)[Y_m][Z_m]
// arguments), behind extern "C" entry points the parity tests load with
// ctypes.  Compiled per grid size: -DY_m=<ny+4> -DZ_m=<nz+4>.
#include "cgli_reference_api.hpp"

typedef double (*A3)[Y_m][Z_m];
static A3 a(double *p) { return reinterpret_cast<A3>(p); }

extern "C" {

int probe_dims(int *out)
{
    out[0] = Y_m;
    out[1] = Z_m;
    return 0;
}

// dir 0/1/2: Flux_X/Y/Z; m: the 9 face metrics of that direction
int probe_flux(void *ctx, int dir, double Ep, double e, double **U, double **scratch, double *J, double *Jf,
               double **m, double **F)
{
    cgli_reference_bind(static_cast<cgli_ctx *>(ctx));
    double **S = scratch;   // MR1..5, ML1..5
    try {
        if (dir == 0)
            Flux_X(0, Ep, e, a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]), a(S[0]), a(S[1]), a(S[2]), a(S[3]),
                   a(S[4]), a(S[5]), a(S[6]), a(S[7]), a(S[8]), a(S[9]), a(J), a(Jf), a(m[0]), a(m[1]), a(m[2]),
                   a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]), a(F[0]), a(F[1]), a(F[2]), a(F[3]),
                   a(F[4]), a(S[0]));
        else if (dir == 1)
            Flux_Y(0, Ep, e, a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]), a(S[0]), a(S[1]), a(S[2]), a(S[3]),
                   a(S[4]), a(S[5]), a(S[6]), a(S[7]), a(S[8]), a(S[9]), a(J), a(Jf), a(m[0]), a(m[1]), a(m[2]),
                   a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]), a(F[0]), a(F[1]), a(F[2]), a(F[3]),
                   a(F[4]), a(S[0]));
        else
            Flux_Z(0, Ep, e, a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]), a(S[0]), a(S[1]), a(S[2]), a(S[3]),
                   a(S[4]), a(S[5]), a(S[6]), a(S[7]), a(S[8]), a(S[9]), a(J), a(Jf), a(m[0]), a(m[1]), a(m[2]),
                   a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]), a(F[0]), a(F[1]), a(F[2]), a(F[3]),
                   a(F[4]), a(S[0]));
    } catch (...) {
        return -1;
    }
    return 0;
}

// mc/mu/mv/mw: 9 metrics each; S: 30 scratch arrays
int probe_viscous(void *ctx, double **U, double **vF, double *J, double **mc, double *Ju, double **mu, double *Jv,
                  double **mv, double *Jw, double **mw, double **S)
{
    cgli_reference_bind(static_cast<cgli_ctx *>(ctx));
    try {
        Viscous_terms(0, a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]), a(vF[0]), a(vF[1]), a(vF[2]), a(vF[3]),
                      a(J), a(mc[0]), a(mc[1]), a(mc[2]), a(mc[3]), a(mc[4]), a(mc[5]), a(mc[6]), a(mc[7]),
                      a(mc[8]), a(Ju), a(mu[0]), a(mu[1]), a(mu[2]), a(mu[3]), a(mu[4]), a(mu[5]), a(mu[6]),
                      a(mu[7]), a(mu[8]), a(Jv), a(mv[0]), a(mv[1]), a(mv[2]), a(mv[3]), a(mv[4]), a(mv[5]),
                      a(mv[6]), a(mv[7]), a(mv[8]), a(Jw), a(mw[0]), a(mw[1]), a(mw[2]), a(mw[3]), a(mw[4]),
                      a(mw[5]), a(mw[6]), a(mw[7]), a(mw[8]), a(S[0]), a(S[1]), a(S[2]), a(S[3]), a(S[4]),
                      a(S[5]), a(S[6]), a(S[7]), a(S[8]), a(S[9]), a(S[10]), a(S[11]), a(S[12]), a(S[13]),
                      a(S[14]), a(S[15]), a(S[16]), a(S[17]), a(S[18]), a(S[19]), a(S[20]), a(S[21]), a(S[22]),
                      a(S[23]), a(S[24]), a(S[25]), a(S[26]), a(S[27]), a(S[28]), a(S[29]));
    } catch (...) {
        return -1;
    }
    return 0;
}

// jacobi = 0: Dual_time_stepping (Nijk/ijk, etdy_v, J_v used), 1: DPLUSGS
int probe_implicit(void *ctx, int jacobi, double deltaT, double e, double *er, int *Nijk, int *ijk, double **U_,
                   double **U, double **Uq, double *J, double **m, double **Fx, double **Fy, double **Fz,
                   double **vF, double **MR, double **NR, double **NL, double **Res, double *etdy_v, double *J_v)
{
    cgli_reference_bind(static_cast<cgli_ctx *>(ctx));
    try {
        if (!jacobi)
            Dual_time_stepping(0, deltaT, e, &er[0], &er[1], &er[2], &er[3], &er[4],
                               reinterpret_cast<int (*)[2]>(Nijk), reinterpret_cast<int (*)[3]>(ijk), a(U_[0]),
                               a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]), a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]),
                               a(Uq[0]), a(Uq[1]), a(Uq[2]), a(Uq[3]), a(Uq[4]), a(J), a(m[0]), a(m[1]), a(m[2]),
                               a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]), a(m[8]), a(Fx[0]), a(Fx[1]), a(Fx[2]),
                               a(Fx[3]), a(Fx[4]), a(Fy[0]), a(Fy[1]), a(Fy[2]), a(Fy[3]), a(Fy[4]), a(Fz[0]),
                               a(Fz[1]), a(Fz[2]), a(Fz[3]), a(Fz[4]), a(vF[0]), a(vF[1]), a(vF[2]), a(vF[3]),
                               a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]), a(NR[0]), a(NR[1]), a(NR[2]),
                               a(NR[3]), a(NR[4]), a(NL[0]), a(NL[1]), a(NL[2]), a(NL[3]), a(NL[4]), a(Res[0]),
                               a(Res[1]), a(Res[2]), a(Res[3]), a(Res[4]), a(etdy_v), a(J_v));
        else
            DPLUSGS(0, deltaT, &er[0], &er[1], &er[2], &er[3], &er[4], e, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]),
                    a(U_[4]), a(U[0]), a(U[1]), a(U[2]), a(U[3]), a(U[4]), a(Uq[0]), a(Uq[1]), a(Uq[2]), a(Uq[3]),
                    a(Uq[4]), a(J), a(m[0]), a(m[1]), a(m[2]), a(m[3]), a(m[4]), a(m[5]), a(m[6]), a(m[7]),
                    a(m[8]), a(Fx[0]), a(Fx[1]), a(Fx[2]), a(Fx[3]), a(Fx[4]), a(Fy[0]), a(Fy[1]), a(Fy[2]),
                    a(Fy[3]), a(Fy[4]), a(Fz[0]), a(Fz[1]), a(Fz[2]), a(Fz[3]), a(Fz[4]), a(vF[0]), a(vF[1]),
                    a(vF[2]), a(vF[3]), a(MR[0]), a(MR[1]), a(MR[2]), a(MR[3]), a(MR[4]), a(NR[0]), a(NR[1]),
                    a(NR[2]), a(NR[3]), a(NR[4]), a(NL[0]), a(NL[1]), a(NL[2]), a(NL[3]), a(NL[4]), a(Res[0]),
                    a(Res[1]), a(Res[2]), a(Res[3]), a(Res[4]));
    } catch (...) {
        return -1;
    }
    return 0;
}

int probe_filter(void *ctx, double crit, double E_high, double E_low, double **U_, double *Y, double *EpX,
                 double *EpY, double *EpZ, double *J)
{
    cgli_reference_bind(static_cast<cgli_ctx *>(ctx));
    try {
        Filter(0, crit, E_high, E_low, a(U_[0]), a(U_[1]), a(U_[2]), a(U_[3]), a(U_[4]), a(Y), a(EpX), a(EpY),
               a(EpZ), a(J));
    } catch (...) {
        return -1;
    }
    return 0;
}

int probe_hyperplane(void *ctx, int *Nijk, int *ijk)
{
    cgli_reference_bind(static_cast<cgli_ctx *>(ctx));
    try {
        Hyperplane(0, reinterpret_cast<int (*)[2]>(Nijk), reinterpret_cast<int (*)[3]>(ijk));
    } catch (...) {
        return -1;
    }
    return 0;
}

}  // extern "C"
