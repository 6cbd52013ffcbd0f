/*
 * oracle_flux.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Restatement of the reference inviscid fluxes:
 *   Flux_X.cpp:92-357   (MUSCL-5 J-weighted face states + preconditioned Roe)
 *   Flux_Y.cpp:94-413   (near-wall order drop, wall-face pressure override)
 *   Flux_Z.cpp:91-359   (order drop at both z ends)
 * Arithmetic is kept in the reference's evaluation order so that results
 * agree to the last bit on an x86-64 build without FMA contraction.
 *
 * The *_scaled variants add the dissipation scaling of the older flux
 * files (Flux_X~~.cpp:552-556, Flux_Y~~.cpp, Flux_Z~~.cpp): the Roe
 * dissipation Fav of every face is multiplied by a factor, 1 (mode 0, the
 * live reference), the scalar Ep (mode 1) or the face's filter-sensor value
 * EpF (mode 2, face-indexed like the fluxes).  Multiplying by 1 is exact,
 * so mode 0 is the unscaled flux bit for bit.
 */
#include "oracle_common.h"

/* Roe-type preconditioned face flux (Flux_X.cpp:251-350; identical in
 * Flux_Y.cpp:295-396 and Flux_Z.cpp:254-354).  (mx,my,mz) are the face
 * metric components, (nx_,ny_,nz_) the vector multiplying deltaP in
 * Fav2..4.  Returns the flux and, for the wall override, iP, P_ and Fav3. */
static void roe_face(const double Lq[5], const double Rq[5], double mx, double my, double mz,
                     double Jf, double e, double K, double nx_, double ny_, double nz_, double ep,
                     double F[5], double *iP_out, double *P_out, double *Fav3_out)
{
    double insqr = 1.0 / sqrt(mx * mx + my * my + mz * mz);

    double rL = Lq[0];
    double uL = Lq[1] / rL, vL = Lq[2] / rL, wL = Lq[3] / rL;
    double UnL = mx * uL + my * vL + mz * wL;
    double qL = uL * uL + vL * vL + wL * wL;
    double pL = (Lq[4] - 0.5 * rL * qL) * (K - 1);
    double cL = K * pL / rL;
    double hL = 0.5 * qL + cL / (K - 1);

    double rR = Rq[0];
    double uR = Rq[1] / rR, vR = Rq[2] / rR, wR = Rq[3] / rR;
    double UnR = mx * uR + my * vR + mz * wR;
    double qR = uR * uR + vR * vR + wR * wR;
    double pR = (Rq[4] - 0.5 * rR * qR) * (K - 1);
    double cR = K * pR / rR;
    double hR = 0.5 * qR + cR / (K - 1);

    double sL = sqrt(rL), sR = sqrt(rR);
    double ss = sL + sR;
    double rho = sqrt(rL * rR);
    double u = (sL * uL + sR * uR) / ss;
    double v = (sL * vL + sR * vR) / ss;
    double w = (sL * wL + sR * wR) / ss;
    double Un = (sL * UnL + sR * UnR) / ss;
    double q = u * u + v * v + w * w;
    double H = (sL * hL + sR * hR) / ss;
    double C = (H - 0.5 * q) * (K - 1);   /* squared sound speed */

    double dU1 = rR - rL;
    double dU2 = rR * uR - rL * uL;
    double dU3 = rR * vR - rL * vL;
    double dU4 = rR * wR - rL * wL;
    double dU5 = (pR / (K - 1) + 0.5 * rR * qR) - (pL / (K - 1) + 0.5 * rL * qL);

    double beta = OMAX(q / C, e);
    double ubar = 0.5 * (1 + beta) * Un * insqr;
    double S = 0.5 * sqrt(4 * beta * C + Un * insqr * Un * insqr * (1 - beta) * (1 - beta));
    double th = q / C;
    double U_p = 0.5 * (1 + th) * Un * insqr;
    double C_p = 0.5 * sqrt(4 * C * th + Un * insqr * Un * insqr * (1 - th) * (1 - th));

    double t1 = (pR - pL) / rho / beta / C;
    double t2 = U_p / S * (UnR - UnL) * insqr;
    double t3 = 0.5 * (1 - beta) * Un * insqr * ubar / S;
    double t4 = 0.5 * (1 - th) * Un * insqr * U_p / S;
    double dUc = (S - t3 - beta * fabs(Un) * insqr) * t1 + t2;
    double dPc = U_p / S * (pR - pL) + (C_p - fabs(Un) * insqr + t4) * rho * (UnR - UnL) * insqr;

    double Fav1 = fabs(Un * insqr) * dU1 + dUc * rho;
    double Fav2 = fabs(Un * insqr) * dU2 + dUc * rho * u + dPc * nx_;
    double Fav3 = fabs(Un * insqr) * dU3 + dUc * rho * v + dPc * ny_;
    double Fav4 = fabs(Un * insqr) * dU4 + dUc * rho * w + dPc * nz_;
    double Fav5 = fabs(Un * insqr) * dU5 + dUc * rho * H + dPc * Un * insqr;

    F[0] = 0.5 * ((rL * UnL + rR * UnR) - ep * Fav1 / insqr) / Jf;
    F[1] = 0.5 * ((rL * uL * UnL + rR * uR * UnR + mx * (pL + pR)) - ep * Fav2 / insqr) / Jf;
    F[2] = 0.5 * ((rL * vL * UnL + rR * vR * UnR + my * (pL + pR)) - ep * Fav3 / insqr) / Jf;
    F[3] = 0.5 * ((rL * wL * UnL + rR * wR * UnR + mz * (pL + pR)) - ep * Fav4 / insqr) / Jf;
    F[4] = 0.5 * ((UnL * (3.5 * pL + 0.5 * rL * qL) + UnR * (3.5 * pR + 0.5 * rR * qR)) - ep * Fav5 / insqr) / Jf;
    if (iP_out) *iP_out = pL;
    if (P_out) *P_out = pR;
    if (Fav3_out) *Fav3_out = Fav3;
}

/* 5-point J-weighted reconstruction (Flux_X.cpp:112-126): left-biased
 * weights (2,-13,47,27,-3) for the state left of face +1/2, and the mirror
 * (-3,27,47,-13,2) for the state right of face -1/2. */
static double rec5L(const double *U, const double *J, long a, long b, long c, long d, long f,
                    double *temp_out)
{
    double temp = 1. / (2 / J[a] - 13 / J[b] + 47 / J[c] + 27 / J[d] - 3 / J[f]);
    (void)temp_out;
    return temp * (2 * U[a] - 13 * U[b] + 47 * U[c] + 27 * U[d] - 3 * U[f]);
}
static double rec5R(const double *U, const double *J, long a, long b, long c, long d, long f)
{
    double temp = 1. / (-3 / J[a] + 27 / J[b] + 47 / J[c] - 13 / J[d] + 2 / J[f]);
    return temp * (-3 * U[a] + 27 * U[b] + 47 * U[c] - 13 * U[d] + 2 * U[f]);
}
/* third-order one-sided pair (Flux_X.cpp:155-170) */
static double rec3L(const double *U, const double *J, long m, long c, long p)
{
    double temp = 1. / (-1 / J[m] + 5 / J[c] + 2 / J[p]);
    return temp * (-U[m] + 5 * U[c] + 2 * U[p]);
}
static double rec3R(const double *U, const double *J, long m, long c, long p)
{
    double temp = 1. / (2 / J[m] + 5 / J[c] - 1 / J[p]);
    return temp * (2 * U[m] + 5 * U[c] - U[p]);
}
/* two-point J-weighted average (Flux_X.cpp:145-151) */
static double rec2(const double *U, const double *J, long a, long b)
{
    double temp = 1. / (1 / J[a] + 1 / J[b]);
    return temp * (U[a] + U[b]);
}

static double face_scale(int mode, double Ep, const double *EpF, long id)
{
    return mode == 2 ? EpF[id] : (mode == 1 ? Ep : 1.0);
}

void orc_flux_x_scaled(const orc_cfg *c, int mode, double Ep, double e, double *const U_[5],
                       double *MR[5], double *ML[5], const double *J, const double *J_u,
                       const double *const met_u[9], double *inFx[5], const double *EpF)
{
    DIMS(c);
    int istart = c->first ? 4 : 2;
    int iend = c->last ? gend : gend + 1;
    #pragma omp parallel for schedule(static)
    for (int i = istart; i <= iend; i++)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++)
                for (int v = 0; v < 5; v++) {
                    ML[v][ID(i, j - 1, k - 1)] = rec5L(U_[v], J, ID(i - 2, j, k), ID(i - 1, j, k), ID(i, j, k),
                                                       ID(i + 1, j, k), ID(i + 2, j, k), 0);
                    MR[v][ID(i - 1, j - 1, k - 1)] = rec5R(U_[v], J, ID(i - 2, j, k), ID(i - 1, j, k), ID(i, j, k),
                                                           ID(i + 1, j, k), ID(i + 2, j, k));
                }
    if (c->first)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++)
                for (int v = 0; v < 5; v++) {
                    ML[v][ID(2, j - 1, k - 1)] = rec2(U_[v], J, ID(2, j, k), ID(3, j, k));
                    ML[v][ID(3, j - 1, k - 1)] = rec3L(U_[v], J, ID(2, j, k), ID(3, j, k), ID(4, j, k));
                    MR[v][ID(2, j - 1, k - 1)] = rec3R(U_[v], J, ID(2, j, k), ID(3, j, k), ID(4, j, k));
                }
    if (c->last)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++)
                for (int v = 0; v < 5; v++) {
                    int i = gend;
                    ML[v][ID(i, j - 1, k - 1)] = rec3L(U_[v], J, ID(i - 1, j, k), ID(i, j, k), ID(i + 1, j, k));
                    MR[v][ID(i - 1, j - 1, k - 1)] = rec3R(U_[v], J, ID(i - 1, j, k), ID(i, j, k), ID(i + 1, j, k));
                    MR[v][ID(i, j - 1, k - 1)] = rec2(U_[v], J, ID(i, j, k), ID(i + 1, j, k));
                }

    #pragma omp parallel for schedule(static)
    for (int i = 2; i <= gend; i++)
        for (int j = 1; j < NY; j++)
            for (int k = 1; k < NZ; k++) {
                long id = ID(i, j, k);
                double L[5], R[5], F[5];
                for (int v = 0; v < 5; v++) { L[v] = ML[v][id]; R[v] = MR[v][id]; }
                double xix = met_u[0][id], xiy = met_u[1][id], xiz = met_u[2][id];
                double insqr = 1.0 / sqrt(xix * xix + xiy * xiy + xiz * xiz);
                roe_face(L, R, xix, xiy, xiz, J_u[id], e, c->K, xix * insqr, xiy * insqr, xiz * insqr,
                         face_scale(mode, Ep, EpF, id), F, 0, 0, 0);
                for (int v = 0; v < 5; v++) inFx[v][id] = F[v];
            }
}

void orc_flux_y_scaled(const orc_cfg *c, int mode, double Ep, double e, double *const U_[5],
                       double *MR[5], double *ML[5], const double *J, const double *J_v,
                       const double *const met_v[9], double *inFy[5], const double *EpF)
{
    DIMS(c);
    int istart = c->first ? 2 : 3;
    int iend = gend;
    #pragma omp parallel for schedule(static)
    for (int i = istart; i <= iend; i++)
        for (int j = 4; j <= NY - 2; j++)
            for (int k = 2; k <= NZ; k++)
                for (int v = 0; v < 5; v++) {
                    ML[v][ID(i - 1, j, k - 1)] = rec5L(U_[v], J, ID(i, j - 2, k), ID(i, j - 1, k), ID(i, j, k),
                                                       ID(i, j + 1, k), ID(i, j + 2, k), 0);
                    MR[v][ID(i - 1, j - 1, k - 1)] = rec5R(U_[v], J, ID(i, j - 2, k), ID(i, j - 1, k), ID(i, j, k),
                                                           ID(i, j + 1, k), ID(i, j + 2, k));
                }
    #pragma omp parallel for schedule(static)
    for (int i = istart; i <= iend; i++)
        for (int k = 2; k <= NZ; k++)
            for (int v = 0; v < 5; v++) {
                const double *U = U_[v];
                int j = 1;
                ML[v][ID(i - 1, j, k - 1)] = rec2(U, J, ID(i, j, k), ID(i, j + 1, k));
                j = 2;
                ML[v][ID(i - 1, j, k - 1)] = rec2(U, J, ID(i, j, k), ID(i, j + 1, k));
                MR[v][ID(i - 1, j - 1, k - 1)] = rec2(U, J, ID(i, j, k), ID(i, j + 1, k));
                j = 3;
                ML[v][ID(i - 1, j, k - 1)] = rec3L(U, J, ID(i, j - 1, k), ID(i, j, k), ID(i, j + 1, k));
                MR[v][ID(i - 1, j - 1, k - 1)] = rec3R(U, J, ID(i, j - 1, k), ID(i, j, k), ID(i, j + 1, k));
                j = NY - 1;
                ML[v][ID(i - 1, j, k - 1)] = rec3L(U, J, ID(i, j - 1, k), ID(i, j, k), ID(i, j + 1, k));
                MR[v][ID(i - 1, j - 1, k - 1)] = rec3R(U, J, ID(i, j - 1, k), ID(i, j, k), ID(i, j + 1, k));
                j = NY;
                ML[v][ID(i - 1, j, k - 1)] = rec2(U, J, ID(i, j, k), ID(i, j - 1, k));
                MR[v][ID(i - 1, j - 1, k - 1)] = rec2(U, J, ID(i, j, k), ID(i, j - 1, k));
                MR[v][ID(i - 1, j, k - 1)] = rec2(U, J, ID(i, j, k), ID(i, j + 1, k));
            }

    #pragma omp parallel for schedule(static)
    for (int i = 2; i <= gend - 1; i++)
        for (int j = 1; j < NYY; j++)
            for (int k = 1; k < NZ; k++) {
                long id = ID(i, j, k);
                double L[5], R[5], F[5], iP, P_, Fav3;
                for (int v = 0; v < 5; v++) { L[v] = ML[v][id]; R[v] = MR[v][id]; }
                double etx = met_v[3][id], ety = met_v[4][id], etz = met_v[5][id];
                double insqr = 1.0 / sqrt(etx * etx + ety * ety + etz * etz);
                double nx_ = 0, ny_ = 0, nz_ = 0;
                if (c->xi_mode) { nx_ = etx * insqr; ny_ = ety * insqr; nz_ = etz * insqr; }
                roe_face(L, R, etx, ety, etz, J_v[id], e, c->K, nx_, ny_, nz_, face_scale(mode, Ep, EpF, id), F,
                         &iP, &P_, &Fav3);
                for (int v = 0; v < 5; v++) inFy[v][id] = F[v];
                if (j == 1 || j == NY) {
                    inFy[0][id] = 0;
                    inFy[1][id] = 0;
                    inFy[2][id] = 0.5 * (ety * (iP + P_) - Ep * Fav3) / J_v[id];
                    inFy[3][id] = 0;
                    inFy[4][id] = 0;
                }
            }
}

void orc_flux_z_scaled(const orc_cfg *c, int mode, double Ep, double e, double *const U_[5],
                       double *MR[5], double *ML[5], const double *J, const double *J_w,
                       const double *const met_w[9], double *inFz[5], const double *EpF)
{
    DIMS(c);
    int istart = c->first ? 2 : 3;
    int iend = gend;
    #pragma omp parallel for schedule(static)
    for (int i = istart; i <= iend; i++)
        for (int j = 2; j <= NY; j++)
            for (int k = 3; k <= NZ - 1; k++)
                for (int v = 0; v < 5; v++) {
                    ML[v][ID(i - 1, j - 1, k)] = rec5L(U_[v], J, ID(i, j, k - 2), ID(i, j, k - 1), ID(i, j, k),
                                                       ID(i, j, k + 1), ID(i, j, k + 2), 0);
                    MR[v][ID(i - 1, j - 1, k - 1)] = rec5R(U_[v], J, ID(i, j, k - 2), ID(i, j, k - 1), ID(i, j, k),
                                                           ID(i, j, k + 1), ID(i, j, k + 2));
                }
    #pragma omp parallel for schedule(static)
    for (int i = istart; i <= iend; i++)
        for (int j = 2; j <= NY; j++)
            for (int v = 0; v < 5; v++) {
                const double *U = U_[v];
                int k = 1;
                ML[v][ID(i - 1, j - 1, k)] = rec2(U, J, ID(i, j, k), ID(i, j, k + 1));
                k = 2;
                ML[v][ID(i - 1, j - 1, k)] = rec3L(U, J, ID(i, j, k - 1), ID(i, j, k), ID(i, j, k + 1));
                MR[v][ID(i - 1, j - 1, k - 1)] = rec3R(U, J, ID(i, j, k - 1), ID(i, j, k), ID(i, j, k + 1));
                k = NZ;
                ML[v][ID(i - 1, j - 1, k)] = rec3L(U, J, ID(i, j, k - 1), ID(i, j, k), ID(i, j, k + 1));
                MR[v][ID(i - 1, j - 1, k - 1)] = rec3R(U, J, ID(i, j, k - 1), ID(i, j, k), ID(i, j, k + 1));
                MR[v][ID(i - 1, j - 1, k)] = rec2(U, J, ID(i, j, k), ID(i, j, k + 1));
            }

    #pragma omp parallel for schedule(static)
    for (int i = 2; i <= gend - 1; i++)
        for (int j = 1; j < NY; j++)
            for (int k = 1; k < NZZ; k++) {
                long id = ID(i, j, k);
                double L[5], R[5], F[5];
                for (int v = 0; v < 5; v++) { L[v] = ML[v][id]; R[v] = MR[v][id]; }
                double ztx = met_w[6][id], zty = met_w[7][id], ztz = met_w[8][id];
                double insqr = 1.0 / (sqrt(ztx * ztx + zty * zty + ztz * ztz));
                double nx_ = 0, ny_ = 0, nz_ = 0;
                if (c->xi_mode) { nx_ = ztx * insqr; ny_ = zty * insqr; nz_ = ztz * insqr; }
                roe_face(L, R, ztx, zty, ztz, J_w[id], e, c->K, nx_, ny_, nz_, face_scale(mode, Ep, EpF, id), F, 0,
                         0, 0);
                for (int v = 0; v < 5; v++) inFz[v][id] = F[v];
            }
}

/* the live reference fluxes: dissipation unscaled */
void orc_flux_x(const orc_cfg *c, double Ep, double e, double *const U_[5], double *MR[5], double *ML[5],
                const double *J, const double *J_u, const double *const met_u[9], double *inFx[5])
{
    orc_flux_x_scaled(c, 0, Ep, e, U_, MR, ML, J, J_u, met_u, inFx, 0);
}

void orc_flux_y(const orc_cfg *c, double Ep, double e, double *const U_[5], double *MR[5], double *ML[5],
                const double *J, const double *J_v, const double *const met_v[9], double *inFy[5])
{
    orc_flux_y_scaled(c, 0, Ep, e, U_, MR, ML, J, J_v, met_v, inFy, 0);
}

void orc_flux_z(const orc_cfg *c, double Ep, double e, double *const U_[5], double *MR[5], double *ML[5],
                const double *J, const double *J_w, const double *const met_w[9], double *inFz[5])
{
    orc_flux_z_scaled(c, 0, Ep, e, U_, MR, ML, J, J_w, met_w, inFz, 0);
}
