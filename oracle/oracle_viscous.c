/*
 * oracle_viscous.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Restatement of Viscous_terms.cpp:183-905: for every cell (i,j,k) with
 * i=3..gend+1, j=2..NYY, k=2..NZZ the viscous/heat flux through its low
 * x face (into LL), low y face (into ML) and low z face (into NL), using
 * second-order face gradients with 4-point cross averages and Sutherland
 * viscosity; then vF = dLL/dxi + dML/deta + dNL/dzeta on interior cells.
 * The kUj*iVj / kVj*iWj / kUj*iWj products of Viscous_terms.cpp:744-746
 * are reproduced as written.
 */
#include "oracle_common.h"

typedef struct { double u, v, w, T; } prim4;

/* Viscous_terms.cpp:260-266: primitive (u,v,w,T) of cell id; U1_ is rho/J */
static prim4 prim(const orc_cfg *c, double *const U_[5], const double *J, long id)
{
    prim4 p;
    double r = U_[0][id];
    p.u = U_[1][id] / r;
    p.v = U_[2][id] / r;
    p.w = U_[3][id] / r;
    double P = (U_[4][id] - 0.5 * r * (p.u * p.u + p.v * p.v + p.w * p.w)) * (c->K - 1) * J[id];
    p.T = P / (r * c->R) / J[id];
    return p;
}

/* metric groups a..f of Viscous_terms.cpp:464-519 */
typedef struct { double a[8], b[8], cc[8], d[11], e[11], f[11]; } mgroups;

static void groups(const double m[9], mgroups *g)
{
    double xix = m[0], xiy = m[1], xiz = m[2], etx = m[3], ety = m[4], etz = m[5], ztx = m[6],
           zty = m[7], ztz = m[8];
    g->a[1] = (4. / 3) * xix * xix + xiy * xiy + xiz * xiz;
    g->a[2] = xix * xix + (4. / 3) * xiy * xiy + xiz * xiz;
    g->a[3] = xix * xix + xiy * xiy + (4. / 3) * xiz * xiz;
    g->a[4] = xix * xix + xiy * xiy + xiz * xiz;
    g->a[5] = (1. / 3) * xix * xiy;
    g->a[6] = (1. / 3) * xiy * xiz;
    g->a[7] = (1. / 3) * xix * xiz;
    g->b[1] = (4. / 3) * etx * etx + ety * ety + etz * etz;
    g->b[2] = etx * etx + (4. / 3) * ety * ety + etz * etz;
    g->b[3] = etx * etx + ety * ety + (4. / 3) * etz * etz;
    g->b[4] = etx * etx + ety * ety + etz * etz;
    g->b[5] = (1. / 3) * etx * ety;
    g->b[6] = (1. / 3) * ety * etz;
    g->b[7] = (1. / 3) * etx * etz;
    g->cc[1] = (4. / 3) * ztx * ztx + zty * zty + ztz * ztz;
    g->cc[2] = ztx * ztx + (4. / 3) * zty * zty + ztz * ztz;
    g->cc[3] = ztx * ztx + zty * zty + (4. / 3) * ztz * ztz;
    g->cc[4] = ztx * ztx + zty * zty + ztz * ztz;
    g->cc[5] = (1. / 3) * ztx * zty;
    g->cc[6] = (1. / 3) * zty * ztz;
    g->cc[7] = (1. / 3) * ztx * ztz;
    g->d[1] = (4. / 3) * xix * etx + xiy * ety + xiz * etz;
    g->d[2] = xix * etx + (4. / 3) * xiy * ety + xiz * etz;
    g->d[3] = xix * etx + xiy * ety + (4. / 3) * xiz * etz;
    g->d[4] = xix * etx + xiy * ety + xiz * etz;
    g->d[5] = xix * ety - (2. / 3) * xiy * etx;
    g->d[6] = xix * etz - (2. / 3) * xiz * etx;
    g->d[7] = xiy * etx - (2. / 3) * xix * ety;
    g->d[8] = xiy * etz - (2. / 3) * xiz * ety;
    g->d[9] = xiz * etx - (2. / 3) * xix * etz;
    g->d[10] = xiz * ety - (2. / 3) * xiy * etz;
    g->e[1] = (4. / 3) * xix * ztx + xiy * zty + xiz * ztz;
    g->e[2] = xix * ztx + (4. / 3) * xiy * zty + xiz * ztz;
    g->e[3] = xix * ztx + xiy * zty + (4. / 3) * xiz * ztz;
    g->e[4] = xix * ztx + xiy * zty + xiz * ztz;
    g->e[5] = xix * zty - (2. / 3) * xiy * ztx;
    g->e[6] = xix * ztz - (2. / 3) * xiz * ztx;
    g->e[7] = xiy * ztx - (2. / 3) * xix * zty;
    g->e[8] = xiy * ztz - (2. / 3) * xiz * zty;
    g->e[9] = xiz * ztx - (2. / 3) * xix * ztz;
    g->e[10] = xiz * zty - (2. / 3) * xiy * ztz;
    g->f[1] = (4. / 3) * etx * ztx + ety * zty + etz * ztz;
    g->f[2] = etx * ztx + (4. / 3) * ety * zty + etz * ztz;
    g->f[3] = etx * ztx + ety * zty + (4. / 3) * etz * ztz;
    g->f[4] = etx * ztx + ety * zty + etz * ztz;
    g->f[5] = etx * zty - (2. / 3) * ety * ztx;
    g->f[6] = etx * ztz - (2. / 3) * etz * ztx;
    g->f[7] = ety * ztx - (2. / 3) * etx * zty;
    g->f[8] = ety * ztz - (2. / 3) * etz * zty;
    g->f[9] = etz * ztx - (2. / 3) * etx * ztz;
    g->f[10] = etz * zty - (2. / 3) * ety * ztz;
}

/* gradient set of one face: first derivatives, derivatives of products and
 * of T along x, y, z (Viscous_terms.cpp:399-450) */
typedef struct {
    double u, v, w, u2, v2, w2, uv, vw, uw, T;
} dset;

/* two-point difference (P - M) * inv */
static dset diff2(prim4 P, prim4 M, double inv)
{
    dset d;
    d.u = (P.u - M.u) * inv;
    d.v = (P.v - M.v) * inv;
    d.w = (P.w - M.w) * inv;
    d.u2 = (P.u * P.u - M.u * M.u) * inv;
    d.v2 = (P.v * P.v - M.v * M.v) * inv;
    d.w2 = (P.w * P.w - M.w * M.w) * inv;
    d.uv = (P.u * P.v - M.u * M.v) * inv;
    d.vw = (P.v * P.w - M.v * M.w) * inv;
    d.uw = (P.u * P.w - M.u * M.w) * inv;
    d.T = (P.T - M.T) * inv;
    return d;
}
/* four-point cross difference (A + B - C - D) * inv */
static dset diff4(prim4 A, prim4 B, prim4 C, prim4 D, double inv)
{
    dset d;
    d.u = (A.u + B.u - C.u - D.u) * inv;
    d.v = (A.v + B.v - C.v - D.v) * inv;
    d.w = (A.w + B.w - C.w - D.w) * inv;
    d.u2 = (A.u * A.u + B.u * B.u - C.u * C.u - D.u * D.u) * inv;
    d.v2 = (A.v * A.v + B.v * B.v - C.v * C.v - D.v * D.v) * inv;
    d.w2 = (A.w * A.w + B.w * B.w - C.w * C.w - D.w * D.w) * inv;
    d.uv = (A.u * A.v + B.u * B.v - C.u * C.v - D.u * D.v) * inv;
    d.vw = (A.v * A.w + B.v * B.w - C.v * C.w - D.v * D.w) * inv;
    d.uw = (A.u * A.w + B.u * B.w - C.u * C.w - D.u * D.w) * inv;
    d.T = (A.T + B.T - C.T - D.T) * inv;
    return d;
}

static double sutherland(const orc_cfg *c, double T)
{
    return c->mu_L * pow((T / c->suth_T), 1.5) * (c->suth_T + c->suth_S) / (T + c->suth_S);
}

void orc_viscous(const orc_cfg *c, double *const U_[5], double *vF[4],
                 const double *J, const double *J_u, const double *const met_u[9],
                 const double *J_v, const double *const met_v[9],
                 const double *J_w, const double *const met_w[9],
                 double *LL[5], double *ML[5], double *NL[5])
{
    DIMS(c);
    const double invXI = 1. / (c->dxi), invET = 1. / (c->deta), invZT = 1. / (c->dzeta);
    const double inv4XI = 1. / (4 * c->dxi), inv4ET = 1. / (4 * c->deta), inv4ZT = 1. / (4 * c->dzeta);
    const double CvK = c->Cv * c->K, Pr = c->Pr_L;

    #pragma omp parallel for schedule(static)
    for (int i = 3; i <= gend + 1; i++)
        for (int j = 2; j <= NYY; j++)
            for (int k = 2; k <= NZZ; k++) {
                prim4 O = prim(c, U_, J, ID(i, j, k));
                prim4 iOj = prim(c, U_, J, ID(i - 1, j + 1, k));
                prim4 iOk = prim(c, U_, J, ID(i - 1, j, k + 1));
                prim4 jOi = prim(c, U_, J, ID(i + 1, j - 1, k));
                prim4 kOi = prim(c, U_, J, ID(i + 1, j, k - 1));
                prim4 jOk = prim(c, U_, J, ID(i, j - 1, k + 1));
                prim4 kOj = prim(c, U_, J, ID(i, j + 1, k - 1));
                prim4 iO = prim(c, U_, J, ID(i - 1, j, k));
                prim4 jO = prim(c, U_, J, ID(i, j - 1, k));
                prim4 kO = prim(c, U_, J, ID(i, j, k - 1));
                prim4 ijO = prim(c, U_, J, ID(i - 1, j - 1, k));
                prim4 jkO = prim(c, U_, J, ID(i, j - 1, k - 1));
                prim4 ikO = prim(c, U_, J, ID(i - 1, j, k - 1));
                prim4 Oi = prim(c, U_, J, ID(i + 1, j, k));
                prim4 Oj = prim(c, U_, J, ID(i, j + 1, k));
                prim4 Ok = prim(c, U_, J, ID(i, j, k + 1));
                long fid = ID(i - 1, j - 1, k - 1);
                long out = ID(i, j, k);
                mgroups g;
                double m[9];

                /* ---- x face i-1/2 (Viscous_terms.cpp:399-546) ---- */
                {
                    dset X = diff2(O, iO, invXI);
                    dset Y = diff4(Oj, iOj, jO, ijO, inv4ET);
                    dset Z = diff4(Ok, iOk, kO, ikO, inv4ZT);
                    for (int q = 0; q < 9; q++) m[q] = met_u[q][fid];
                    groups(m, &g);
                    double Ux = 0.5 * (O.u + iO.u), Vx = 0.5 * (O.v + iO.v), Wx = 0.5 * (O.w + iO.w),
                           Tx = 0.5 * (O.T + iO.T);
                    double mu = sutherland(c, Tx);
                    double Jf = J_u[fid];
                    LL[1][out] = mu * (g.a[1] * X.u + g.a[5] * X.v + g.a[7] * X.w + g.d[1] * Y.u + g.d[7] * Y.v +
                                       g.d[9] * Y.w + g.e[1] * Z.u + g.e[7] * Z.v + g.e[9] * Z.w) / Jf;
                    LL[2][out] = mu * (g.a[5] * X.u + g.a[2] * X.v + g.a[6] * X.w + g.d[5] * Y.u + g.d[2] * Y.v +
                                       g.d[10] * Y.w + g.e[5] * Z.u + g.e[2] * Z.v + g.e[10] * Z.w) / Jf;
                    LL[3][out] = mu * (g.a[7] * X.u + g.a[6] * X.v + g.a[3] * X.w + g.d[6] * Y.u + g.d[8] * Y.v +
                                       g.d[3] * Y.w + g.e[6] * Z.u + g.e[8] * Z.v + g.e[3] * Z.w) / Jf;
                    LL[4][out] = mu * (0.5 * g.a[1] * X.u2 + 0.5 * g.a[2] * X.v2 + 0.5 * g.a[3] * X.w2 +
                                       g.a[5] * X.uv + g.a[6] * X.vw + g.a[7] * X.uw + CvK * g.a[4] * X.T / (Pr) +
                                       0.5 * g.d[1] * Y.u2 + 0.5 * g.d[2] * Y.v2 + 0.5 * g.d[3] * Y.w2 +
                                       g.d[5] * Vx * Y.u + g.d[6] * Wx * Y.u + g.d[7] * Ux * Y.v + g.d[8] * Wx * Y.v +
                                       g.d[9] * Ux * Y.w + g.d[10] * Vx * Y.w + CvK * g.d[4] * Y.T / (Pr) +
                                       0.5 * g.e[1] * Z.u2 + 0.5 * g.e[2] * Z.v2 + 0.5 * g.e[3] * Z.w2 +
                                       g.e[5] * Vx * Z.u + g.e[6] * Wx * Z.u + g.e[7] * Ux * Z.v + g.e[8] * Wx * Z.v +
                                       g.e[9] * Ux * Z.w + g.e[10] * Vx * Z.w + CvK * g.e[4] * Z.T / (Pr)) / Jf;
                    (void)Tx;
                }
                /* ---- y face j-1/2 (Viscous_terms.cpp:563-700) ---- */
                {
                    dset X = diff4(Oi, jOi, iO, ijO, inv4XI);
                    dset Y = diff2(O, jO, invET);
                    dset Z = diff4(Ok, jOk, kO, jkO, inv4ZT);
                    for (int q = 0; q < 9; q++) m[q] = met_v[q][fid];
                    groups(m, &g);
                    double Uy = 0.5 * (O.u + jO.u), Vy = 0.5 * (O.v + jO.v), Wy = 0.5 * (O.w + jO.w),
                           Ty = 0.5 * (O.T + jO.T);
                    double mu = sutherland(c, Ty);
                    double Jf = J_v[fid];
                    ML[1][out] = mu * (g.d[1] * X.u + g.d[5] * X.v + g.d[6] * X.w + g.b[1] * Y.u + g.b[5] * Y.v +
                                       g.b[7] * Y.w + g.f[1] * Z.u + g.f[7] * Z.v + g.f[9] * Z.w) / Jf;
                    ML[2][out] = mu * (g.d[7] * X.u + g.d[2] * X.v + g.d[8] * X.w + g.b[5] * Y.u + g.b[2] * Y.v +
                                       g.b[6] * Y.w + g.f[5] * Z.u + g.f[2] * Z.v + g.f[10] * Z.w) / Jf;
                    ML[3][out] = mu * (g.d[9] * X.u + g.d[10] * X.v + g.d[3] * X.w + g.b[7] * Y.u + g.b[6] * Y.v +
                                       g.b[3] * Y.w + g.f[6] * Z.u + g.f[8] * Z.v + g.f[3] * Z.w) / Jf;
                    ML[4][out] = mu * (0.5 * g.b[1] * Y.u2 + 0.5 * g.b[2] * Y.v2 + 0.5 * g.b[3] * Y.w2 +
                                       g.b[5] * Y.uv + g.b[6] * Y.vw + g.b[7] * Y.uw + CvK * g.b[4] * Y.T / (Pr) +
                                       0.5 * g.d[1] * X.u2 + 0.5 * g.d[2] * X.v2 + 0.5 * g.d[3] * X.w2 +
                                       g.d[5] * Uy * X.v + g.d[6] * Uy * X.w + g.d[7] * Vy * X.u + g.d[8] * Vy * X.w +
                                       g.d[9] * Wy * X.u + g.d[10] * Wy * X.v + CvK * g.d[4] * X.T / (Pr) +
                                       0.5 * g.f[1] * Z.u2 + 0.5 * g.f[2] * Z.v2 + 0.5 * g.f[3] * Z.w2 +
                                       g.f[5] * Vy * Z.u + g.f[6] * Wy * Z.u + g.f[7] * Uy * Z.v + g.f[8] * Wy * Z.v +
                                       g.f[9] * Uy * Z.w + g.f[10] * Vy * Z.w + CvK * g.f[4] * Z.T / (Pr)) / Jf;
                    (void)Ty;
                }
                /* ---- z face k-1/2 (Viscous_terms.cpp:714-861) ---- */
                {
                    dset X = diff4(Oi, kOi, iO, ikO, inv4XI);
                    dset Y = diff4(Oj, kOj, jO, jkO, inv4ET);
                    /* Viscous_terms.cpp:744-746 multiply kUj by iVj / iWj */
                    Y.uv = (Oj.u * Oj.v + kOj.u * iOj.v - jO.u * jO.v - jkO.u * jkO.v) * inv4ET;
                    Y.vw = (Oj.v * Oj.w + kOj.v * iOj.w - jO.v * jO.w - jkO.v * jkO.w) * inv4ET;
                    Y.uw = (Oj.u * Oj.w + kOj.u * iOj.w - jO.u * jO.w - jkO.u * jkO.w) * inv4ET;
                    dset Z = diff2(O, kO, invZT);
                    for (int q = 0; q < 9; q++) m[q] = met_w[q][fid];
                    groups(m, &g);
                    double Uz = 0.5 * (O.u + kO.u), Vz = 0.5 * (O.v + kO.v), Wz = 0.5 * (O.w + kO.w),
                           Tz = 0.5 * (O.T + kO.T);
                    double mu = sutherland(c, Tz);
                    double Jf = J_w[fid];
                    NL[1][out] = mu * (g.e[1] * X.u + g.e[5] * X.v + g.e[6] * X.w + g.f[1] * Y.u + g.f[5] * Y.v +
                                       g.f[6] * Y.w + g.cc[1] * Z.u + g.cc[5] * Z.v + g.cc[7] * Z.w) / Jf;
                    NL[2][out] = mu * (g.e[7] * X.u + g.e[2] * X.v + g.e[8] * X.w + g.f[7] * Y.u + g.f[2] * Y.v +
                                       g.f[8] * Y.w + g.cc[5] * Z.u + g.cc[2] * Z.v + g.cc[6] * Z.w) / Jf;
                    NL[3][out] = mu * (g.e[9] * X.u + g.e[10] * X.v + g.e[3] * X.w + g.f[9] * Y.u + g.f[10] * Y.v +
                                       g.f[3] * Y.w + g.cc[7] * Z.u + g.cc[6] * Z.v + g.cc[3] * Z.w) / Jf;
                    NL[4][out] = mu * (0.5 * g.cc[1] * Z.u2 + 0.5 * g.cc[2] * Z.v2 + 0.5 * g.cc[3] * Z.w2 +
                                       g.cc[5] * Z.uv + g.cc[6] * Z.vw + g.cc[7] * Z.uw + CvK * g.cc[4] * Z.T / (Pr) +
                                       0.5 * g.e[1] * X.u2 + 0.5 * g.e[2] * X.v2 + 0.5 * g.e[3] * X.w2 +
                                       g.e[5] * Uz * X.v + g.e[6] * Uz * X.w + g.e[7] * Vz * X.u + g.e[8] * Vz * X.w +
                                       g.e[9] * Wz * X.u + g.e[10] * Wz * X.v + CvK * g.e[4] * X.T / (Pr) +
                                       0.5 * g.f[1] * Y.u2 + 0.5 * g.f[2] * Y.v2 + 0.5 * g.f[3] * Y.w2 +
                                       g.f[5] * Uz * Y.v + g.f[6] * Uz * Y.w + g.f[7] * Vz * Y.u + g.f[8] * Vz * Y.w +
                                       g.f[9] * Wz * Y.u + g.f[10] * Wz * Y.v + CvK * g.f[4] * Y.T / (Pr)) / Jf;
                    (void)Tz;
                }
            }

    #pragma omp parallel for schedule(static)
    for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++)
                for (int v = 1; v < 5; v++)
                    vF[v - 1][ID(i, j, k)] = (LL[v][ID(i + 1, j, k)] - LL[v][ID(i, j, k)]) * invXI +
                                             (ML[v][ID(i, j + 1, k)] - ML[v][ID(i, j, k)]) * invET +
                                             (NL[v][ID(i, j, k + 1)] - NL[v][ID(i, j, k)]) * invZT;
}
