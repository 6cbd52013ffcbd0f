/*
 * oracle_filter.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Restatement of the multi-scale filter sensor, Filter.cpp:12-1720.
 *
 * For each of two scales (h: weights 1/8 (1,6,1); 2h: 1/4 (1,2,1)) the
 * velocity-carrying components U1..U4 are passed six times through a
 * separable filter G = Gz Gy Gx (Filter.cpp:143-742 for h, 825-1417 for
 * 2h); the y pass weights the neighbours by the one-sided spacings of
 * Y_point and by J (Filter.cpp:199, 875).  The six levels are combined as
 * 6G - 15G^2 + 20G^3 - 15G^4 + 6G^5 - G^6, i.e. 1-(1-G)^6
 * (Filter.cpp:769-773, 1444), the squared velocity difference to the
 * unfiltered state gives E_h and E_2h (796-800, 1471-1475), their ghost
 * cells copy the adjacent interior (1480-1548), and the face sensors are
 * E_high where (E_h+E_h')/(E_2h+E_2h') >= Roe_criterion, else E_low
 * (1560-1645).
 *
 * Index ranges follow the reference's rank logic with `first`/`last` in
 * place of myid==0 / myid==nproc-1.  The GG halo exchange after the second
 * level (Filter.cpp:433-466, 1106-1141) is a no-op at a first/last end
 * (MPI_PROC_NULL); with x_periodic the slab is its own neighbour on both
 * sides, and the x-ghost E values (never set by the reference on interior
 * ranks) are the periodic images; on an interior slab that is not
 * periodic they stay zero (parity unpinned there: the reference reads
 * memory it never wrote).  U5 is filtered by the reference but
 * never used; it is skipped here.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_common.h"

/* x range of level l (0..5) of one scale: Filter.cpp:149-153 (levels 0, 3),
 * 253-257 (levels 1, 4), 348-349 (levels 2, 5) */
static void level_range(const orc_cfg *c, int lvl, int *lo, int *hi)
{
    const int gend = c->nx + 2;
    const int w = (lvl == 0 || lvl == 3) ? 2 : ((lvl == 1 || lvl == 4) ? 1 : 0);
    *lo = c->first ? 3 : 3 - w;
    *hi = c->last ? gend : gend + w;
}

/* one level: in -> Gx -> Gy -> out over x range lo..hi, j=2..NY, k=2..NZ */
static void level(const orc_cfg *c, int twoh, int lo, int hi, double *const in[4], double *Gx[4],
                  double *Gy[4], double *out[4], const double *Y, const double *J)
{
    DIMS(c);
    for (int v = 0; v < 4; v++) {
        const double *u = in[v];
        double *gx = Gx[v];
        for (int i = lo; i <= hi; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++) {
                    const long id = ID(i, j, k);
                    if (!twoh)
                        gx[id] = 0.125 * (u[ID(i - 1, j, k)] + 6 * u[id] + u[ID(i + 1, j, k)]);
                    else
                        gx[id] = 0.25 * (u[ID(i - 1, j, k)] + 2 * u[id] + u[ID(i + 1, j, k)]);
                }
    }
    for (int v = 0; v < 4; v++) {
        const double *gx = Gx[v];
        double *gy = Gy[v];
        for (int i = lo; i <= hi; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++) {
                    const long id = ID(i, j, k), jm = ID(i, j - 1, k), jp = ID(i, j + 1, k);
                    const double a = Y[jp] - Y[jm];
                    const double b = Y[jp] - Y[id];
                    const double cc = Y[id] - Y[jm];
                    if (!twoh)
                        gy[id] = 0.125 * (2 * b / a * gx[jm] * J[jm] + 6 * gx[id] * J[id] +
                                          2 * cc / a * gx[jp] * J[jp]) / J[id];
                    else
                        gy[id] = 0.5 * (b / a * gx[jm] * J[jm] + gx[id] * J[id] + cc / a * gx[jp] * J[jp]) / J[id];
                }
    }
    for (int v = 0; v < 4; v++) {
        const double *gy = Gy[v];
        double *o = out[v];
        for (int i = lo; i <= hi; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++) {
                    const long id = ID(i, j, k);
                    if (!twoh)
                        o[id] = 0.125 * (gy[id - 1] + 6 * gy[id] + gy[id + 1]);
                    else
                        o[id] = 0.25 * (gy[id - 1] + 2 * gy[id] + gy[id + 1]);
                }
    }
}

/* the GG halo exchange, single slab: periodic both ways, else nothing */
static void exchange(const orc_cfg *c, double *G[4])
{
    DIMS(c);
    if (!c->x_periodic) return;
    for (int v = 0; v < 4; v++)
        for (int q = 0; q < 3; q++)
            for (int j = 1; j <= NYY; j++)
                for (int k = 1; k <= NZZ; k++) {
                    G[v][ID(gend + 1 + q, j, k)] = G[v][ID(3 + q, j, k)];
                    G[v][ID(q, j, k)] = G[v][ID(gend - 2 + q, j, k)];
                }
}

static void scale(const orc_cfg *c, int twoh, double *const U_[5], const double *Y, const double *J,
                  double *L[6][4], double *Gx[4], double *Gy[4], double *E)
{
    DIMS(c);
    for (int l = 0; l < 6; l++) {
        int lo, hi;
        level_range(c, l, &lo, &hi);
        double *in[4];
        for (int v = 0; v < 4; v++) in[v] = l == 0 ? U_[v] : L[l - 1][v];
        level(c, twoh, lo, hi, in, Gx, Gy, L[l], Y, J);
        if (l == 2) exchange(c, L[2]);
    }
    double *fU[4];
    for (int v = 0; v < 4; v++) fU[v] = Gx[v]; /* Gx is free again */
    for (int v = 0; v < 4; v++)
        for (int i = 3; i <= gend; i++)
            for (int j = 2; j <= NY; j++)
                for (int k = 2; k <= NZ; k++) {
                    const long id = ID(i, j, k);
                    fU[v][id] = 6 * L[0][v][id] - 15 * L[1][v][id] + 20 * L[2][v][id] - 15 * L[3][v][id] +
                                6 * L[4][v][id] - L[5][v][id];
                }
    for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++) {
                const long id = ID(i, j, k);
                const double Ue = U_[1][id] / U_[0][id] - fU[1][id] / fU[0][id];
                const double Ve = U_[2][id] / U_[0][id] - fU[2][id] / fU[0][id];
                const double We = U_[3][id] / U_[0][id] - fU[3][id] / fU[0][id];
                E[id] = Ue * Ue + Ve * Ve + We * We;
            }
}

/* Filter.cpp:1480-1548 (and the periodic images for x_periodic) */
static void e_ghosts(const orc_cfg *c, double *E)
{
    DIMS(c);
    if (c->first || c->x_periodic)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++) E[ID(2, j, k)] = E[ID(c->x_periodic && !c->first ? gend : 3, j, k)];
    if (c->last || c->x_periodic)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++) E[ID(gend + 1, j, k)] = E[ID(c->x_periodic && !c->last ? 3 : gend, j, k)];
    for (int i = 3; i <= gend; i++)
        for (int k = 2; k <= NZ; k++) {
            E[ID(i, 1, k)] = E[ID(i, 2, k)];
            E[ID(i, NYY, k)] = E[ID(i, NY, k)];
        }
    for (int i = 3; i <= gend; i++)
        for (int j = 2; j <= NY; j++) {
            E[ID(i, j, 1)] = E[ID(i, j, 2)];
            E[ID(i, j, NZZ)] = E[ID(i, j, NZ)];
        }
}

void orc_filter(const orc_cfg *c, double Roe_criterion, double E_high, double E_low,
                double *const U_[5], const double *Y_point,
                double *EpX, double *EpY, double *EpZ, const double *J)
{
    DIMS(c);
    const long N = (long)(c->nx + 6) * Ym * Zm;
    /* 6 levels + Gx + Gy per variable, E_h, E_2h */
    double *pool = (double *)calloc((size_t)(8 * 4 + 2) * N, sizeof(double));
    if (!pool) return;
    double *L[6][4], *Gx[4], *Gy[4];
    for (int l = 0; l < 6; l++)
        for (int v = 0; v < 4; v++) L[l][v] = pool + (l * 4 + v) * N;
    for (int v = 0; v < 4; v++) {
        Gx[v] = pool + (24 + v) * N;
        Gy[v] = pool + (28 + v) * N;
    }
    double *Eh = pool + 32 * N, *E2h = pool + 33 * N;
    /* every work array starts as a copy of U_ (Filter.cpp:117-134) */
    for (int v = 0; v < 4; v++) {
        double *dst[8] = {L[0][v], L[1][v], L[2][v], L[3][v], L[4][v], L[5][v], Gx[v], Gy[v]};
        for (int i = 0; i <= gend + 3; i++)
            for (int j = 1; j <= NYY; j++)
                for (int k = 1; k <= NZZ; k++)
                    for (int q = 0; q < 8; q++) dst[q][ID(i, j, k)] = U_[v][ID(i, j, k)];
    }
    scale(c, 0, U_, Y_point, J, L, Gx, Gy, Eh);
    scale(c, 1, U_, Y_point, J, L, Gx, Gy, E2h);
    e_ghosts(c, Eh);
    e_ghosts(c, E2h);
    for (int i = 2; i <= gend; i++)
        for (int j = 1; j < NY; j++)
            for (int k = 1; k < NZ; k++) {
                const double Emax = (Eh[ID(i, j + 1, k + 1)] + Eh[ID(i + 1, j + 1, k + 1)]) /
                                    (E2h[ID(i, j + 1, k + 1)] + E2h[ID(i + 1, j + 1, k + 1)]);
                EpX[ID(i, j, k)] = (Emax >= Roe_criterion) ? E_high : E_low;
            }
    for (int i = 2; i <= gend - 1; i++)
        for (int j = 1; j < NYY; j++)
            for (int k = 1; k < NZ; k++) {
                const double Emax = (Eh[ID(i + 1, j, k + 1)] + Eh[ID(i + 1, j + 1, k + 1)]) /
                                    (E2h[ID(i + 1, j, k + 1)] + E2h[ID(i + 1, j + 1, k + 1)]);
                EpY[ID(i, j, k)] = (Emax >= Roe_criterion) ? E_high : E_low;
            }
    for (int i = 2; i <= gend - 1; i++)
        for (int j = 1; j < NY; j++)
            for (int k = 1; k < NZZ; k++) {
                const double Emax = (Eh[ID(i + 1, j + 1, k)] + Eh[ID(i + 1, j + 1, k + 1)]) /
                                    (E2h[ID(i + 1, j + 1, k)] + E2h[ID(i + 1, j + 1, k + 1)]);
                EpZ[ID(i, j, k)] = (Emax >= Roe_criterion) ? E_high : E_low;
            }
    free(pool);
}

/* the intermediate E fields, for tests of the sensor's inputs */
void orc_filter_energy(const orc_cfg *c, double *const U_[5], const double *Y_point, const double *J,
                       double *Eh_out, double *E2h_out)
{
    DIMS(c);
    const long N = (long)(c->nx + 6) * Ym * Zm;
    double *pool = (double *)calloc((size_t)(8 * 4) * N, sizeof(double));
    if (!pool) return;
    double *L[6][4], *Gx[4], *Gy[4];
    for (int l = 0; l < 6; l++)
        for (int v = 0; v < 4; v++) L[l][v] = pool + (l * 4 + v) * N;
    for (int v = 0; v < 4; v++) {
        Gx[v] = pool + (24 + v) * N;
        Gy[v] = pool + (28 + v) * N;
    }
    for (int v = 0; v < 4; v++) {
        double *dst[8] = {L[0][v], L[1][v], L[2][v], L[3][v], L[4][v], L[5][v], Gx[v], Gy[v]};
        for (int i = 0; i <= gend + 3; i++)
            for (int j = 1; j <= NYY; j++)
                for (int k = 1; k <= NZZ; k++)
                    for (int q = 0; q < 8; q++) dst[q][ID(i, j, k)] = U_[v][ID(i, j, k)];
    }
    scale(c, 0, U_, Y_point, J, L, Gx, Gy, Eh_out);
    scale(c, 1, U_, Y_point, J, L, Gx, Gy, E2h_out);
    e_ghosts(c, Eh_out);
    e_ghosts(c, E2h_out);
    free(pool);
}
