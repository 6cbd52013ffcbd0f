/*
 * oracle_dts.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Restatement of the implicit inner iteration Dual_time_stepping.cpp:18-1792
 * (ghost zeroing, sponge sources, BDF2 RHS transformed by Gamma^-1, spectral
 * radius diagonal, symmetric LU-SGS sweeps on i+j+k hyperplanes, primitive
 * update and L2 norm), the Jacobi variant DPLUSGS.cpp:18-1238 and the
 * hyperplane ordering Hyperplane.cpp:38-78.  One slab, no MPI: the
 * reference's MPI_Sendrecv calls are with MPI_PROC_NULL at nproc=1 and its
 * MPI_Allreduce is the identity.
 */
#include <string.h>
#include "oracle_common.h"

typedef struct { double rho, u, v, w, VV, P, C, T, H, beta; } cprim;

/* Dual_time_stepping.cpp:759-769: primitives of cell id from Q/J */
static cprim cellp(const orc_cfg *c, double *const U_[5], const double *J, long id, double e)
{
    cprim q;
    double K = c->K;
    q.rho = U_[0][id] * J[id];
    q.u = U_[1][id] / U_[0][id];
    q.v = U_[2][id] / U_[0][id];
    q.w = U_[3][id] / U_[0][id];
    q.VV = q.u * q.u + q.v * q.v + q.w * q.w;
    q.P = (U_[4][id] * J[id] - 0.5 * q.rho * q.VV) * (K - 1);
    q.C = K * q.P / q.rho;
    q.T = q.P / q.rho;
    q.H = 0.5 * q.VV + q.C / (K - 1);
    q.beta = OMAX(q.VV / q.C, e);
    return q;
}

/* split-flux coefficients of one neighbour (Dual_time_stepping.cpp:1054-1170,
 * 1384-1524).  late52 selects the B/C ordering 0.5*beta*T*(K-1)*m of the
 * x52..x54 entries; the A entries use 0.5*m*beta*T*(K-1). */
typedef struct {
    double p11, p12, p13, p14, p21, p22, p31, p33, p41, p44, p51, p52, p53, p54, p55, m11, m22;
} splitc;

static void split(const cprim *q, double mx, double my, double mz, double K, int late52, splitc *s)
{
    double b = q->beta;
    double Un = mx * q->u + my * q->v + mz * q->w;
    double S = sqrt(Un * Un * (b - 1) * (b - 1) + 4 * b * q->C * (mx * mx + my * my + mz * mz));
    double lam = ((b + 1) * fabs(Un) + S) / 2;
    double d = fabs(lam);
    s->p11 = 0.5 * (b * Un + d);
    s->p12 = 0.5 * q->P * K * b * mx;
    s->p13 = 0.5 * q->P * K * b * my;
    s->p14 = 0.5 * q->P * K * b * mz;
    s->p21 = 0.5 * mx / q->rho;
    s->p22 = 0.5 * (Un + d);
    s->p31 = 0.5 * my / q->rho;
    s->p33 = 0.5 * (Un + d);
    s->p41 = 0.5 * mz / q->rho;
    s->p44 = 0.5 * (Un + d);
    s->p51 = 0.5 * Un * (K - 1) * (b - 1) / q->rho / K;
    if (!late52) {
        s->p52 = 0.5 * mx * b * q->T * (K - 1);
        s->p53 = 0.5 * my * b * q->T * (K - 1);
        s->p54 = 0.5 * mz * b * q->T * (K - 1);
    } else {
        s->p52 = 0.5 * b * q->T * (K - 1) * mx;
        s->p53 = 0.5 * b * q->T * (K - 1) * my;
        s->p54 = 0.5 * b * q->T * (K - 1) * mz;
    }
    s->p55 = 0.5 * (Un + d);
    s->m11 = 0.5 * (b * Un - d);
    s->m22 = 0.5 * (Un - d);
}

/* Dual_time_stepping.cpp:856-895: scalar diagonal with infinite pseudo step */
static void diag(const orc_cfg *c, const cprim *q, const double m[9], double deltaT,
                 double *d11, double *d22, double *d51)
{
    double K = c->K, b = q->beta;
    double U = m[0] * q->u + m[1] * q->v + m[2] * q->w;
    double V = m[3] * q->u + m[4] * q->v + m[5] * q->w;
    double W = m[6] * q->u + m[7] * q->v + m[8] * q->w;
    double Sx = sqrt(U * U * (b - 1) * (b - 1) + 4 * b * q->C * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]));
    double Sy = sqrt(V * V * (b - 1) * (b - 1) + 4 * b * q->C * (m[3] * m[3] + m[4] * m[4] + m[5] * m[5]));
    double Sz = sqrt(W * W * (b - 1) * (b - 1) + 4 * b * q->C * (m[6] * m[6] + m[7] * m[7] + m[8] * m[8]));
    double lX = ((b + 1) * fabs(U) + Sx) / 2;
    double lY = ((b + 1) * fabs(V) + Sy) / 2;
    double lZ = ((b + 1) * fabs(W) + Sz) / 2;
    double lambda = lX / c->dxi + lY / c->deta + lZ / c->dzeta;
    *d11 = 2 * deltaT / (3 * b + 2 * lambda * deltaT);
    *d22 = 2 * deltaT / (3 + 2 * lambda * deltaT);
    double t1 = 6 * (K - 1) * (b - 1) * deltaT;
    double t2 = K * (3 + 2 * lambda * deltaT) * (3 * b + 2 * lambda * deltaT) * q->rho;
    *d51 = -t1 / t2;
}

static void load9(double *const met[9], long id, double m[9])
{
    for (int q = 0; q < 9; q++) m[q] = met[q][id];
}

int orc_hyperplane_count(const orc_cfg *c) { return c->nx * c->ny * c->nz; }

void orc_hyperplane(const orc_cfg *c, int *Nijk, int *ijk)
{
    DIMS(c);
    int ii = 0;
    for (int m = 7; m <= gend + NY + NZ; m++) {
        int mi = 0;
        for (int i = 3; i <= gend; i++)
            for (int j = 2; j < NYY; j++)
                for (int k = 2; k < NZZ; k++)
                    if (m == i + j + k) {
                        mi++;
                        ii++;
                        ijk[3 * ii + 0] = i;
                        ijk[3 * ii + 1] = j;
                        ijk[3 * ii + 2] = k;
                    }
        Nijk[2 * m + 0] = ii;
        Nijk[2 * m + 1] = mi;
    }
}

/* Dual_time_stepping.cpp:198-323 (and DPLUSGS.cpp:185-298) */
static void zero_increment_ghosts(const orc_cfg *c, double *NR[5], double *NL[5])
{
    DIMS(c);
    for (int v = 0; v < 5; v++) {
        if (c->first || c->x_periodic)
            for (int j = 1; j <= NYY; j++)
                for (int k = 1; k <= NZZ; k++) NL[v][ID(2, j, k)] = NR[v][ID(2, j, k)] = 0;
        if (c->last || c->x_periodic)
            for (int j = 1; j <= NYY; j++)
                for (int k = 1; k <= NZZ; k++) NL[v][ID(gend + 1, j, k)] = NR[v][ID(gend + 1, j, k)] = 0;
        for (int i = 2; i <= gend + 1; i++)
            for (int k = 1; k <= NZZ; k++) {
                NL[v][ID(i, 1, k)] = NR[v][ID(i, 1, k)] = 0;
                NL[v][ID(i, NYY, k)] = NR[v][ID(i, NYY, k)] = 0;
            }
        for (int i = 2; i <= gend + 1; i++)
            for (int j = 1; j <= NYY; j++) {
                NL[v][ID(i, j, 1)] = NR[v][ID(i, j, 1)] = 0;
                NL[v][ID(i, j, NZZ)] = NR[v][ID(i, j, NZZ)] = 0;
            }
    }
}

/* Dual_time_stepping.cpp:503-664: outlet target velocity and the inlet /
 * outlet absorbing-zone sources written into Residual */
static void sponge(const orc_cfg *c, double e, double *const U_[5], const double *J,
                   const double *etdy, double *Res[5])
{
    DIMS(c);
    const double K = c->K;
    double z1 = 0.0;
    double DZ = 1. / (c->high * c->dzeta * c->nz);
    if (c->first)
        for (int j = 2; j <= NYY; j++)
            for (int k = 2; k <= NZZ; k++) {
                double U = U_[1][ID(3, j, k)] / U_[0][ID(3, j, k)];
                z1 = z1 + U * c->deta / etdy[ID(3, j, k)] * c->dzeta;
            }
    double targetU = z1 * DZ;
    const double U_in_0 = 1.0, Sigma_in_0 = 0.035, U_out_0 = 1.0, Sigma_out_0 = 1.25;
    const int Xg = c->nx_global;
    for (int zone = 0; zone < 2; zone++) {
        int istart, iend;
        if (zone == 0) {
            if (!(c->gstart < c->nx_inlet)) continue;
            istart = 3;
            iend = (c->gend0 < c->nx_inlet) ? gend : c->nx_inlet - c->gstart + 2;
        } else {
            if (!(c->gend0 > (Xg - c->nx_outlet - 1))) continue;
            istart = (c->gstart > (Xg - c->nx_outlet - 1)) ? 3 : (Xg - c->nx_outlet) - c->gstart + 2;
            iend = gend;
        }
        for (int i = istart; i <= iend; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++) {
                    long id = ID(i, j, k), im = ID(i - 1, j, k);
                    double rho = U_[0][id] * J[id];
                    double u = U_[1][id] / U_[0][id];
                    double v = U_[2][id] / U_[0][id];
                    double w = U_[3][id] / U_[0][id];
                    double VV = u * u + v * v + w * w;
                    double P = (U_[4][id] * J[id] - 0.5 * rho * VV) * (K - 1);
                    double C = K * P / rho;
                    double beta = OMAX(VV / C, e);
                    double C_plan = 0.5 * sqrt(u * u * (beta - 1) * (beta - 1) + 4 * beta * C);
                    double s1, sg;
                    double t[5];
                    if (zone == 0) {
                        s1 = pow((((c->nx_inlet - c->gstart - 0.5) * 1.0) / c->nx_inlet), 3.0) * U_in_0 * C_plan;
                        sg = Sigma_in_0 * s1 / c->high;
                        t[0] = rho - c->rho_ref;
                        t[1] = rho * u;
                        t[2] = rho * v;
                        t[3] = rho * w;
                        t[4] = U_[4][id] * J[id] - c->E_ref;
                    } else {
                        int nxr = Xg + 1;
                        s1 = pow((((i + c->gstart - 0.5 - (nxr - c->nx_outlet)) * 1.0) / c->nx_outlet), 3.0) *
                             U_out_0 * C_plan;
                        sg = Sigma_out_0 * s1 / c->high;
                        t[0] = rho - c->rho_ref;
                        t[1] = rho * u - c->rho_ref * targetU;
                        t[2] = rho * v;
                        t[3] = rho * w;
                        t[4] = U_[4][id] * J[id] - c->E_ref - 0.5 * c->rho_ref * targetU * targetU;
                    }
                    for (int q = 0; q < 5; q++)
                        Res[q][id] = -s1 * (U_[q][id] * J[id] - U_[q][im] * J[im]) / c->dxi - sg * t[q];
                }
    }
}

/* Dual_time_stepping.cpp:737-908 / DPLUSGS.cpp:334-483: MR = Gamma^-1 Rk,
 * NL = NR = D MR.  with_res adds Residual/J first (DTS); DPLUSGS omits it. */
static void rhs_diag(const orc_cfg *c, double deltaT, double e, int with_res,
                     double *const U_[5], double *const U[5], double *const Uq[5],
                     const double *J, double *const met[9],
                     double *const inFx[5], double *const inFy[5], double *const inFz[5],
                     double *const vF[4], double *MR[5], double *NR[5], double *NL[5],
                     double *const Res[5])
{
    DIMS(c);
    const double K = c->K;
#pragma omp parallel for schedule(static)
    for (int i = 3; i <= gend; i++)
        for (int j = 2; j < NYY; j++)
            for (int k = 2; k < NZZ; k++) {
                long id = ID(i, j, k);
                cprim q = cellp(c, U_, J, id, e);
                double rho = q.rho, u = q.u, v = q.v, w = q.w, VV = q.VV, T = q.T, H = q.H, beta = q.beta;
                double temp = rho * K;
                double d[5][5];
                d[0][0] = -(H * (K - 1) + VV - K * (T + VV)) * beta;
                d[0][1] = -(K - 1) * u * beta;
                d[0][2] = -(K - 1) * v * beta;
                d[0][3] = -(K - 1) * w * beta;
                d[0][4] = (K - 1) * beta;
                d[1][0] = -u / rho; d[1][1] = 1 / rho; d[1][2] = 0; d[1][3] = 0; d[1][4] = 0;
                d[2][0] = -v / rho; d[2][1] = 0; d[2][2] = 1 / rho; d[2][3] = 0; d[2][4] = 0;
                d[3][0] = -w / rho; d[3][1] = 0; d[3][2] = 0; d[3][3] = 1 / rho; d[3][4] = 0;
                d[4][0] = (K - 1) * (VV + K * T * beta + (K - 1) * VV * beta + H * (-1 + beta - K * beta)) / temp;
                d[4][1] = -(K - 1) * u * (1 + (K - 1) * beta) / temp;
                d[4][2] = -(K - 1) * v * (1 + (K - 1) * beta) / temp;
                d[4][3] = -(K - 1) * w * (1 + (K - 1) * beta) / temp;
                d[4][4] = (K - 1) * (1 + (K - 1) * beta) / temp;
                double Rk[5];
                double f = -(rho - c->rho_ref) * c->grav + c->f_const;
                for (int v5 = 0; v5 < 5; v5++) {
                    double Rp = -(3 * U_[v5][id] - 4 * U[v5][id] + Uq[v5][id]) / (2 * deltaT);
                    double Rf = -((inFx[v5][ID(i, j - 1, k - 1)] - inFx[v5][ID(i - 1, j - 1, k - 1)]) / c->dxi +
                                  (inFy[v5][ID(i - 1, j, k - 1)] - inFy[v5][ID(i - 1, j - 1, k - 1)]) / c->deta +
                                  (inFz[v5][ID(i - 1, j - 1, k)] - inFz[v5][ID(i - 1, j - 1, k - 1)]) / c->dzeta);
                    double r = with_res ? Res[v5][id] / J[id] + Rp + Rf : Rp + Rf;
                    if (v5 >= 1) r = r + vF[v5 - 1][id];
                    if (v5 == 1) r = r + f / J[id];
                    if (v5 == 4) r = r + f * u / J[id];
                    Rk[v5] = r;
                }
                for (int a = 0; a < 5; a++)
                    MR[a][id] = (d[a][0] * Rk[0] + d[a][1] * Rk[1] + d[a][2] * Rk[2] + d[a][3] * Rk[3] + d[a][4] * Rk[4]);
                double m[9], d11, d22, d51;
                load9(met, id, m);
                diag(c, &q, m, deltaT, &d11, &d22, &d51);
                NL[0][id] = d11 * MR[0][id];
                NL[1][id] = d22 * MR[1][id];
                NL[2][id] = d22 * MR[2][id];
                NL[3][id] = d22 * MR[3][id];
                NL[4][id] = d51 * MR[0][id] + d22 * MR[4][id];
                for (int a = 0; a < 5; a++) NR[a][id] = NL[a][id];
            }
}

/* lower-neighbour product L(x) of the forward sweep for one cell
 * (Dual_time_stepping.cpp:1172-1232); X are the neighbour increments */
static void lower_sum(const splitc *A, const splitc *B, const splitc *Cc, double *const X[5],
                      long ia, long jb, long kc, const orc_cfg *c, double LL[5])
{
    LL[0] = (-A->p11 * X[0][ia] - A->p12 * X[1][ia] - A->p13 * X[2][ia] - A->p14 * X[3][ia]) / c->dxi +
            (-B->p11 * X[0][jb] - B->p12 * X[1][jb] - B->p13 * X[2][jb] - B->p14 * X[3][jb]) / c->deta +
            (-Cc->p11 * X[0][kc] - Cc->p12 * X[1][kc] - Cc->p13 * X[2][kc] - Cc->p14 * X[3][kc]) / c->dzeta;
    LL[1] = (-A->p21 * X[0][ia] - A->p22 * X[1][ia]) / c->dxi +
            (-B->p21 * X[0][jb] - B->p22 * X[1][jb]) / c->deta +
            (-Cc->p21 * X[0][kc] - Cc->p22 * X[1][kc]) / c->dzeta;
    LL[2] = (-A->p31 * X[0][ia] - A->p33 * X[2][ia]) / c->dxi +
            (-B->p31 * X[0][jb] - B->p33 * X[2][jb]) / c->deta +
            (-Cc->p31 * X[0][kc] - Cc->p33 * X[2][kc]) / c->dzeta;
    LL[3] = (-A->p41 * X[0][ia] - A->p44 * X[3][ia]) / c->dxi +
            (-B->p41 * X[0][jb] - B->p44 * X[3][jb]) / c->deta +
            (-Cc->p41 * X[0][kc] - Cc->p44 * X[3][kc]) / c->dzeta;
    LL[4] = (-A->p51 * X[0][ia] - A->p52 * X[1][ia] - A->p53 * X[2][ia] - A->p54 * X[3][ia] - A->p55 * X[4][ia]) / c->dxi +
            (-B->p51 * X[0][jb] - B->p52 * X[1][jb] - B->p53 * X[2][jb] - B->p54 * X[3][jb] - B->p55 * X[4][jb]) / c->deta +
            (-Cc->p51 * X[0][kc] - Cc->p52 * X[1][kc] - Cc->p53 * X[2][kc] - Cc->p54 * X[3][kc] - Cc->p55 * X[4][kc]) / c->dzeta;
}

/* upper-neighbour product U(x) of the backward sweep
 * (Dual_time_stepping.cpp:1529-1590): minus-split diagonal entries */
static void upper_sum(const splitc *A, const splitc *B, const splitc *Cc, double *const X[5],
                      long ia, long jb, long kc, const orc_cfg *c, double UU[5])
{
    UU[0] = (+A->m11 * X[0][ia] + A->p12 * X[1][ia] + A->p13 * X[2][ia] + A->p14 * X[3][ia]) / c->dxi +
            (+B->m11 * X[0][jb] + B->p12 * X[1][jb] + B->p13 * X[2][jb] + B->p14 * X[3][jb]) / c->deta +
            (+Cc->m11 * X[0][kc] + Cc->p12 * X[1][kc] + Cc->p13 * X[2][kc] + Cc->p14 * X[3][kc]) / c->dzeta;
    UU[1] = (+A->p21 * X[0][ia] + A->m22 * X[1][ia]) / c->dxi +
            (+B->p21 * X[0][jb] + B->m22 * X[1][jb]) / c->deta +
            (+Cc->p21 * X[0][kc] + Cc->m22 * X[1][kc]) / c->dzeta;
    UU[2] = (+A->p31 * X[0][ia] + A->m22 * X[2][ia]) / c->dxi +
            (+B->p31 * X[0][jb] + B->m22 * X[2][jb]) / c->deta +
            (+Cc->p31 * X[0][kc] + Cc->m22 * X[2][kc]) / c->dzeta;
    UU[3] = (+A->p41 * X[0][ia] + A->m22 * X[3][ia]) / c->dxi +
            (+B->p41 * X[0][jb] + B->m22 * X[3][jb]) / c->deta +
            (+Cc->p41 * X[0][kc] + Cc->m22 * X[3][kc]) / c->dzeta;
    UU[4] = (+A->p51 * X[0][ia] + A->p52 * X[1][ia] + A->p53 * X[2][ia] + A->p54 * X[3][ia] + A->m22 * X[4][ia]) / c->dxi +
            (+B->p51 * X[0][jb] + B->p52 * X[1][jb] + B->p53 * X[2][jb] + B->p54 * X[3][jb] + B->m22 * X[4][jb]) / c->deta +
            (+Cc->p51 * X[0][kc] + Cc->p52 * X[1][kc] + Cc->p53 * X[2][kc] + Cc->p54 * X[3][kc] + Cc->m22 * X[4][kc]) / c->dzeta;
}

static void neighbour_split(const orc_cfg *c, double *const U_[5], const double *J, double *const met[9],
                            long id, int dir, double e, splitc *s)
{
    cprim q = cellp(c, U_, J, id, e);
    split(&q, met[3 * dir + 0][id], met[3 * dir + 1][id], met[3 * dir + 2][id], c->K, dir != 0, s);
}

/* Dual_time_stepping.cpp:1702-1782: primitive update and the norm */
static void update_norm(const orc_cfg *c, double *U_[5], const double *J, double *const NR[5],
                        double *Res[5], int dplusgs, double er[5])
{
    DIMS(c);
    const double K = c->K;
#pragma omp parallel for schedule(static)
    for (int i = 3; i <= gend; i++)
        for (int j = 2; j < NYY; j++)
            for (int k = 2; k < NZZ; k++) {
                long id = ID(i, j, k);
                double rhoold = U_[0][id] * J[id];
                double Uold = U_[1][id] / U_[0][id];
                double Vold = U_[2][id] / U_[0][id];
                double Wold = U_[3][id] / U_[0][id];
                double VVold = Uold * Uold + Vold * Vold + Wold * Wold;
                double Pold = (U_[4][id] * J[id] - 0.5 * rhoold * VVold) * (K - 1);
                double Told = Pold / rhoold;
                double P = Pold + NR[0][id] * J[id];
                double U = Uold + NR[1][id] * J[id];
                double V = Vold + NR[2][id] * J[id];
                double W = Wold + NR[3][id] * J[id];
                double T = Told + NR[4][id] * J[id];
                double rho = P / T;
                U_[0][id] = rho / J[id];
                U_[1][id] = rho * U / J[id];
                U_[2][id] = rho * V / J[id];
                U_[3][id] = rho * W / J[id];
                U_[4][id] = (P / (K - 1) + 0.5 * rho * (U * U + V * V + W * W)) / J[id];
                Res[0][id] = dplusgs ? (P - Pold) * (P - Pold) : (rho - rhoold) * (rho - rhoold);
                Res[1][id] = (U - Uold) * (U - Uold);
                Res[2][id] = (V - Vold) * (V - Vold);
                Res[3][id] = (W - Wold) * (W - Wold);
                Res[4][id] = (T - Told) * (T - Told);
            }
    double DN = 1. / (double)((long)c->nx_global * c->ny * c->nz);
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++)
                for (int v = 0; v < 5; v++) {
                    double r = Res[v][ID(i, j, k)];
                    s[v] = s[v] + (dplusgs ? r * r : r);
                }
    for (int v = 0; v < 5; v++) er[v] = sqrt(s[v]) * DN;
}

void orc_dual_time_stepping(const orc_cfg *c, double deltaT, double e, double er[5],
                            const int *Nijk, const int *ijk,
                            double *U_[5], double *const U[5], double *const Uq[5],
                            const double *J, const double *const met[9],
                            double *const inFx[5], double *const inFy[5], double *const inFz[5],
                            double *const vF[4], double *MR[5], double *NR[5], double *NL[5],
                            double *Res[5], const double *etdy_v, const double *J_v, int sweeps)
{
    DIMS(c);
    (void)etdy_v;
    (void)J_v;
    double *const *mt = (double *const *)met;
    zero_increment_ghosts(c, NR, NL);
    long total = (long)(c->nx + 6) * Ym * Zm;
    for (int v = 0; v < 5; v++) memset(Res[v], 0, sizeof(double) * total);
    if (c->sponge) sponge(c, e, U_, J, met[4], Res);
    rhs_diag(c, deltaT, e, 1, U_, U, Uq, J, mt, inFx, inFy, inFz, vF, MR, NR, NL, Res);

    for (int sweep = 1; sweep <= sweeps; sweep++) {
        for (int m = 7; m <= gend + NY + NZ; m++) {
#pragma omp parallel for schedule(static)
            for (int mi = Nijk[2 * m] - Nijk[2 * m + 1] + 1; mi <= Nijk[2 * m]; mi++) {
                int i = ijk[3 * mi], j = ijk[3 * mi + 1], k = ijk[3 * mi + 2];
                long id = ID(i, j, k), ia = ID(i - 1, j, k), jb = ID(i, j - 1, k), kc = ID(i, j, k - 1);
                splitc A, B, Cc;
                neighbour_split(c, (double *const *)U_, J, mt, ia, 0, e, &A);
                neighbour_split(c, (double *const *)U_, J, mt, jb, 1, e, &B);
                neighbour_split(c, (double *const *)U_, J, mt, kc, 2, e, &Cc);
                double LL[5];
                lower_sum(&A, &B, &Cc, NL, ia, jb, kc, c, LL);
                cprim q = cellp(c, (double *const *)U_, J, id, e);
                double m9[9], d11, d22, d51;
                load9(mt, id, m9);
                diag(c, &q, m9, deltaT, &d11, &d22, &d51);
                NL[0][id] = d11 * (MR[0][id] - LL[0]);
                NL[1][id] = d22 * (MR[1][id] - LL[1]);
                NL[2][id] = d22 * (MR[2][id] - LL[2]);
                NL[3][id] = d22 * (MR[3][id] - LL[3]);
                NL[4][id] = d51 * (MR[0][id] - LL[0]) + d22 * (MR[4][id] - LL[4]);
            }
        }
        for (int m = gend + NY + NZ; m >= 7; m--) {
#pragma omp parallel for schedule(static)
            for (int mi = Nijk[2 * m] - Nijk[2 * m + 1] + 1; mi <= Nijk[2 * m]; mi++) {
                int i = ijk[3 * mi], j = ijk[3 * mi + 1], k = ijk[3 * mi + 2];
                long id = ID(i, j, k), ia = ID(i + 1, j, k), jb = ID(i, j + 1, k), kc = ID(i, j, k + 1);
                splitc A, B, Cc;
                neighbour_split(c, (double *const *)U_, J, mt, ia, 0, e, &A);
                neighbour_split(c, (double *const *)U_, J, mt, jb, 1, e, &B);
                neighbour_split(c, (double *const *)U_, J, mt, kc, 2, e, &Cc);
                double UU[5];
                upper_sum(&A, &B, &Cc, NR, ia, jb, kc, c, UU);
                cprim q = cellp(c, (double *const *)U_, J, id, e);
                double m9[9], d11, d22, d51;
                load9(mt, id, m9);
                diag(c, &q, m9, deltaT, &d11, &d22, &d51);
                NR[0][id] = NL[0][id] - d11 * UU[0];
                NR[1][id] = NL[1][id] - d22 * UU[1];
                NR[2][id] = NL[2][id] - d22 * UU[2];
                NR[3][id] = NL[3][id] - d22 * UU[3];
                NR[4][id] = NL[4][id] - (d51 * UU[0] + d22 * UU[4]);
            }
        }
        for (int i = 3; i <= gend; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++)
                    for (int v = 0; v < 5; v++) NL[v][ID(i, j, k)] = NR[v][ID(i, j, k)];
    }
    update_norm(c, U_, J, NR, Res, 0, er);
}

/* DPLUSGS.cpp:484-1150: Jacobi sweeps using the previous sweep's increments
 * on both sides, then NL <- NR */
void orc_dplusgs(const orc_cfg *c, double deltaT, double er[5], double e,
                 double *U_[5], double *const U[5], double *const Uq[5],
                 const double *J, const double *const met[9],
                 double *const inFx[5], double *const inFy[5], double *const inFz[5],
                 double *const vF[4], double *MR[5], double *NR[5], double *NL[5],
                 double *Res[5], int sweeps)
{
    DIMS(c);
    double *const *mt = (double *const *)met;
    orc_cfg cc = *c;
    cc.grav = 0.0;  /* DPLUSGS.cpp applies the constant force f of prm.h:11 */
    cc.f_const = c->f_dplusgs;
    zero_increment_ghosts(c, NR, NL);
    rhs_diag(&cc, deltaT, e, 0, U_, U, Uq, J, mt, inFx, inFy, inFz, vF, MR, NR, NL, Res);
    for (int sweep = 1; sweep <= sweeps; sweep++) {
#pragma omp parallel for schedule(static)
        for (int i = 3; i <= gend; i++)
            for (int j = 2; j < NYY; j++)
                for (int k = 2; k < NZZ; k++) {
                    long id = ID(i, j, k);
                    splitc A, B, Cc, Au, Bu, Cu;
                    long ia = ID(i - 1, j, k), jb = ID(i, j - 1, k), kc = ID(i, j, k - 1);
                    long iu = ID(i + 1, j, k), ju = ID(i, j + 1, k), ku = ID(i, j, k + 1);
                    neighbour_split(c, (double *const *)U_, J, mt, ia, 0, e, &A);
                    neighbour_split(c, (double *const *)U_, J, mt, jb, 1, e, &B);
                    neighbour_split(c, (double *const *)U_, J, mt, kc, 2, e, &Cc);
                    double LL[5], UU[5];
                    lower_sum(&A, &B, &Cc, NL, ia, jb, kc, c, LL);
                    neighbour_split(c, (double *const *)U_, J, mt, iu, 0, e, &Au);
                    neighbour_split(c, (double *const *)U_, J, mt, ju, 1, e, &Bu);
                    neighbour_split(c, (double *const *)U_, J, mt, ku, 2, e, &Cu);
                    upper_sum(&Au, &Bu, &Cu, NL, iu, ju, ku, c, UU);
                    cprim q = cellp(c, (double *const *)U_, J, id, e);
                    double m9[9], d11, d22, d51;
                    load9(mt, id, m9);
                    diag(c, &q, m9, deltaT, &d11, &d22, &d51);
                    NR[0][id] = d11 * (MR[0][id] - LL[0] - UU[0]);
                    NR[1][id] = d22 * (MR[1][id] - LL[1] - UU[1]);
                    NR[2][id] = d22 * (MR[2][id] - LL[2] - UU[2]);
                    NR[3][id] = d22 * (MR[3][id] - LL[3] - UU[3]);
                    NR[4][id] = d51 * (MR[0][id] - LL[0] - UU[0]) + d22 * (MR[4][id] - LL[4] - UU[4]);
                }
        for (int i = 3; i <= gend; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++)
                    for (int v = 0; v < 5; v++) NL[v][ID(i, j, k)] = NR[v][ID(i, j, k)];
    }
    update_norm(c, U_, J, NR, Res, 1, er);
}
