/*
 * oracle_stats.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Wall-normal profiles of the channel: the quantities Statistic.cpp:130-164
 * accumulates (U, V, W, T and the raw second moments UU, VV, WW, TT, UV, UT,
 * VT, with T = p/(rho R) from the conserved state, Statistic.cpp:137-143),
 * averaged over the homogeneous directions x and z of the interior:
 * out[q*ny + (j-2)] for j = 2..NY, q in that order.
 */
#include "oracle_common.h"

void orc_profiles(const orc_cfg *c, double *const U_[5], const double *J, double *out)
{
    DIMS(c);
    const int ny = c->ny;
    const double K = c->K, R = c->R;
    const double inv = 1.0 / ((double)c->nx * (double)c->nz);
    for (int j = 2; j <= NY; j++) {
        double s[11] = {0};
        for (int i = 3; i <= gend; i++)
            for (int k = 2; k <= NZ; k++) {
                const long id = ID(i, j, k);
                const double rho = U_[0][id];
                const double U = U_[1][id] / rho, V = U_[2][id] / rho, W = U_[3][id] / rho;
                const double VV = U * U + V * V + W * W;
                const double P = (U_[4][id] - 0.5 * rho * VV) * (K - 1) * J[id];
                const double T = P / (rho * R) / J[id];
                const double q[11] = {U, V, W, T, U * U, V * V, W * W, T * T, U * V, U * T, V * T};
                for (int a = 0; a < 11; a++) s[a] += q[a];
            }
        for (int a = 0; a < 11; a++) out[a * ny + (j - 2)] = s[a] * inv;
    }
}
