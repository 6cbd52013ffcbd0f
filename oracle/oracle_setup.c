/*
 * oracle_setup.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Reference constants, the deterministic input generator, the analytic
 * metric routines of Jacobian.cpp / Jacobian_u.cpp / Jacobian_v.cpp /
 * Jacobian_w.cpp and the ghost-cell fills of Y_boundary_condition.cpp and
 * Z_boundary_condition.cpp.
 */
#include <limits.h>
#include "oracle_common.h"

void orc_cfg_reference(orc_cfg *c, int nx, int ny, int nz)
{
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->first = 1; c->last = 1;
    c->gstart = 0; c->gend0 = nx - 1; c->nx_global = nx;
    c->nx_inlet = 100; c->nx_outlet = 150;         /* Resolution.h:1-2 */
    c->sponge = 1; c->xi_mode = 0; c->x_periodic = 0; c->precision = 0;
    c->dxi = 0.001; c->deta = 0.005; c->dzeta = 0.002;  /* Resolution.h:19-21 */
    c->high = 0.05;                                 /* prm.h:16 */
    c->K = 1.4; c->R = 287;                         /* prm.h:7-8 */
    c->Cv = 717.5; c->mu_L = 0.0000185; c->Pr_L = 0.71;  /* Viscous_terms.h:3-5 */
    c->suth_T = 298.0592; c->suth_S = 110.;         /* Viscous_terms.cpp:527 */
    c->rho_ref = 1.1842; c->grav = 9.81; c->f_const = 0.0;  /* Dual_time_stepping.cpp:831 */
    c->f_dplusgs = 0.5993;                          /* prm.h:11 */
    c->E_ref = 253250.0;                            /* Dual_time_stepping.cpp:584 */
}

uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double orc_unit(uint64_t *state) { return (double)(orc_splitmix64(state) >> 11) * 0x1.0p-53; }

/* --- Jacobian.cpp:42-243 ------------------------------------------------ */
static void tanh_eta(const orc_cfg *c, double jpos, double *etdy, double *ydet)
{
    double gamma = 2.8;
    double high = c->high;
    double temp = (0.5 * high) * (1 - 1. / tanh(gamma) * tanh(gamma * (1 - 2 * jpos * c->deta)));
    double temp1 = (1 - temp / (0.5 * high)) * tanh(gamma);
    double temp2 = 1. / cosh(gamma * (1 - 2 * jpos * c->deta));
    *etdy = tanh(gamma) / (2 * (0.5 * high) * gamma * (1 - temp1 * temp1));
    *ydet = high * gamma * 1. / tanh(gamma) * temp2 * temp2;
}

static void set_orthogonal(double *J, double *m[9], long id, double etdy, double ydet)
{
    m[0][id] = 1; m[1][id] = 0; m[2][id] = 0;
    m[3][id] = 0; m[4][id] = etdy; m[5][id] = 0;
    m[6][id] = 0; m[7][id] = 0; m[8][id] = 1;
    J[id] = 1. / (1 * ydet * 1);
}

void orc_jacobian(const orc_cfg *c, double *J, double *met[9])
{
    DIMS(c);
    int istart = c->first ? 3 : 0;
    int iend = c->last ? gend : gend + 3;
    for (int i = istart; i <= iend; i++)
        for (int j = 2; j <= NY; j++)
            for (int k = 2; k <= NZ; k++) {
                double etdy, ydet;
                tanh_eta(c, j - 1.5, &etdy, &ydet);
                set_orthogonal(J, met, ID(i, j, k), etdy, ydet);
            }
    /* ghost copies; the reference copies ztdx three times and sets
     * ztdy/ztdz of the j and k ghosts from ztdx (Jacobian.cpp:123-125,
     * 181-182, 224-225) */
    int src[9] = {0, 1, 2, 3, 4, 5, 6, 6, 6};
    if (c->first)
        for (int j = 1; j <= NYY; j++)
            for (int k = 1; k <= NZZ; k++) {
                for (int q = 0; q < 7; q++) met[q][ID(2, j, k)] = met[q][ID(3, j, k)];
                J[ID(2, j, k)] = J[ID(3, j, k)];
            }
    if (c->last)
        for (int j = 1; j <= NYY; j++)
            for (int k = 1; k <= NZZ; k++) {
                for (int q = 0; q < 7; q++) met[q][ID(gend + 1, j, k)] = met[q][ID(gend, j, k)];
                J[ID(gend + 1, j, k)] = J[ID(gend, j, k)];
            }
    istart = c->first ? 2 : 0;
    iend = c->last ? gend + 1 : gend + 3;
    for (int i = istart; i <= iend; i++)
        for (int k = 2; k <= NZ; k++) {
            for (int q = 0; q < 9; q++) met[q][ID(i, 1, k)] = met[src[q]][ID(i, 2, k)];
            for (int q = 0; q < 9; q++) met[q][ID(i, NYY, k)] = met[src[q]][ID(i, NY, k)];
            J[ID(i, 1, k)] = J[ID(i, 2, k)];
            J[ID(i, NYY, k)] = J[ID(i, NY, k)];
        }
    for (int i = istart; i <= iend; i++)
        for (int j = 1; j <= NYY; j++) {
            for (int q = 0; q < 9; q++) met[q][ID(i, j, 1)] = met[src[q]][ID(i, j, 2)];
            for (int q = 0; q < 9; q++) met[q][ID(i, j, NZZ)] = met[src[q]][ID(i, j, NZ)];
            J[ID(i, j, 1)] = J[ID(i, j, 2)];
            J[ID(i, j, NZZ)] = J[ID(i, j, NZ)];
        }
}

/* Jacobian_u.cpp:52-105: x faces, eta at (j-0.5) */
void orc_jacobian_u(const orc_cfg *c, double *J_u, double *met_u[9])
{
    DIMS(c);
    for (int i = 2; i <= gend; i++)
        for (int j = 1; j < NY; j++)
            for (int k = 1; k < NZ; k++) {
                double etdy, ydet;
                tanh_eta(c, j - 0.5, &etdy, &ydet);
                set_orthogonal(J_u, met_u, ID(i, j, k), etdy, ydet);
            }
}

/* Jacobian_v.cpp:58-151: y faces, eta at (j-1); wall faces take the ghost
 * cell-centre metrics */
void orc_jacobian_v(const orc_cfg *c, double *J_v, double *met_v[9],
                    const double *J, double *const met[9])
{
    DIMS(c);
    for (int i = 2; i <= gend - 1; i++)
        for (int j = 1; j <= NY; j++)
            for (int k = 1; k < NZ; k++) {
                double etdy, ydet;
                tanh_eta(c, j - 1.0, &etdy, &ydet);
                set_orthogonal(J_v, met_v, ID(i, j, k), etdy, ydet);
            }
    for (int i = 2; i <= gend - 1; i++)
        for (int k = 1; k < NZ; k++) {
            for (int q = 3; q < 6; q++) {
                met_v[q][ID(i, 1, k)] = met[q][ID(i, 1, k)];
                met_v[q][ID(i, NY, k)] = met[q][ID(i, NYY, k)];
            }
            J_v[ID(i, NY, k)] = J[ID(i, NYY, k)];
            J_v[ID(i, 1, k)] = J[ID(i, 1, k)];
        }
}

/* Jacobian_w.cpp:52-105: z faces, eta at (j-0.5) */
void orc_jacobian_w(const orc_cfg *c, double *J_w, double *met_w[9])
{
    DIMS(c);
    for (int i = 2; i <= gend - 1; i++)
        for (int j = 1; j < NY; j++)
            for (int k = 1; k <= NZ; k++) {
                double etdy, ydet;
                tanh_eta(c, j - 0.5, &etdy, &ydet);
                set_orthogonal(J_w, met_w, ID(i, j, k), etdy, ydet);
            }
}

/* Y_boundary_condition.cpp:98-182: zero-gradient copy of the wall ghosts,
 * then inside (heated_lo, heated_hi) of the global index the isothermal
 * no-slip mirror with the pressure of the first cell. */
void orc_bc_y(const orc_cfg *c, double *U_[5], const double *J, double T_low, double T_high,
              int heated_lo, int heated_hi)
{
    DIMS(c);
    const double K = c->K, R = c->R;
    for (int i = 3; i <= gend; i++) {
        for (int k = 2; k <= NZ; k++)
            for (int v = 0; v < 5; v++) {
                U_[v][ID(i, 1, k)] = U_[v][ID(i, 2, k)];
                U_[v][ID(i, NYY, k)] = U_[v][ID(i, NY, k)];
            }
        if (i + c->gstart > heated_lo && i + c->gstart < heated_hi)
            for (int k = 2; k <= NZ; k++)
                for (int side = 0; side < 2; side++) {
                    int jin = side ? NY : 2, jg = side ? NYY : 1;
                    long in = ID(i, jin, k), g = ID(i, jg, k);
                    double rho = U_[0][in] * J[in];
                    double U = U_[1][in] / U_[0][in];
                    double V = U_[2][in] / U_[0][in];
                    double W = U_[3][in] / U_[0][in];
                    double VV = U * U + V * V + W * W;
                    double P = (U_[4][in] * J[in] - 0.5 * rho * VV) * (K - 1);
                    double T = side ? T_high : T_low;
                    rho = P / R / T;
                    U_[0][g] = rho / J[g];
                    U_[1][g] = -rho * U / J[g];
                    U_[2][g] = -rho * V / J[g];
                    U_[3][g] = -rho * W / J[g];
                    U_[4][g] = (P / (K - 1) + 0.5 * rho * VV) / J[g];
                }
    }
}

void orc_bc_z(const orc_cfg *c, double *U_[5])
{
    DIMS(c);
    int istart = c->first ? 2 : 0;
    int iend = c->last ? gend + 1 : gend + 3;
    for (int i = istart; i <= iend; i++)
        for (int j = 0; j <= NYYY; j++)
            for (int v = 0; v < 5; v++) {
                U_[v][ID(i, j, 1)] = U_[v][ID(i, j, NZ)];
                U_[v][ID(i, j, NZZ)] = U_[v][ID(i, j, 2)];
            }
}

/* One periodic slab: the three ghost planes at each x end copy the
 * interior planes at the other end, every j and k (the single-slab form of
 * the main.cpp:167-200 halo). */
void orc_bc_x_periodic(const orc_cfg *c, double *U_[5])
{
    DIMS(c);
    for (int g = 0; g < 3; g++)
        for (int j = 0; j <= NYYY; j++)
            for (int k = 0; k <= NZZZ; k++)
                for (int v = 0; v < 5; v++) {
                    U_[v][ID(g, j, k)] = U_[v][ID(gend - 2 + g, j, k)];
                    U_[v][ID(gend + 1 + g, j, k)] = U_[v][ID(3 + g, j, k)];
                }
}

/* --- seeded inputs ------------------------------------------------------ */
void orc_fill(double *a, long n, uint64_t seed, double lo, double hi)
{
    uint64_t s = seed;
    for (long i = 0; i < n; i++) a[i] = lo + (hi - lo) * orc_unit(&s);
}

/* Q/J state from primitive draws on every entry of the padded array:
 * rho in [r0,r1], u,v,w in [-ua,ua], p in [p0,p1] (each its own stream) */
void orc_fill_state(long n, uint64_t seed, const double *J, double K, double r0, double r1,
                    double ua, double p0, double p1, double *U_[5])
{
    uint64_t s[5];
    for (int q = 0; q < 5; q++) s[q] = seed * 8 + (uint64_t)q;
    for (long i = 0; i < n; i++) {
        double rho = r0 + (r1 - r0) * orc_unit(&s[0]);
        double u = -ua + 2 * ua * orc_unit(&s[1]);
        double v = -ua + 2 * ua * orc_unit(&s[2]);
        double w = -ua + 2 * ua * orc_unit(&s[3]);
        double p = p0 + (p1 - p0) * orc_unit(&s[4]);
        U_[0][i] = rho / J[i];
        U_[1][i] = rho * u / J[i];
        U_[2][i] = rho * v / J[i];
        U_[3][i] = rho * w / J[i];
        U_[4][i] = (p / (K - 1) + 0.5 * rho * (u * u + v * v + w * w)) / J[i];
    }
}
