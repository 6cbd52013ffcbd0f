// channel.cpp -- the compressible turbulent channel workload and the
// physical-step driver (the inner-iteration sequence of main.cpp:157-386).
//
// Grid: x = xi + a sin(2 pi y) sin(2 pi zeta / Lz), y = -tanh(beta(1-2 eta)) /
// tanh(beta), z = zeta, with xi in [0,Lx), eta in [0,1], zeta in [0,Lz).
// The mapping does not depend on xi, so every metric is a function of
// (j,k) only: one (j,k) plane is computed on the host in FP64 and
// replicated along i on the device.  Cell centres sit at eta = (j-1.5)
// deta (Jacobian.cpp:66), x faces at (j-0.5) (Jacobian_u.cpp:73), y faces at
// (j-1) (Jacobian_v.cpp:93), z faces at (j-0.5) with zeta at the face.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "cgli_internal.hpp"

using namespace cgli;

void cgli::channel_drop_graphs(ChannelState &cs)
{
    for (auto &q : cs.graphs)
        if (q.g) cgli_graph_destroy(q.g);
    cs.graphs.clear();
}

namespace {

const double PI = 3.14159265358979323846;

struct Mapping {
    double Lz, beta, a;
    double y(double eta) const { return -std::tanh(beta * (1 - 2 * eta)) / std::tanh(beta); }
    double y_eta(double eta) const
    {
        const double t = std::tanh(beta * (1 - 2 * eta));
        return 2 * beta * (1 - t * t) / std::tanh(beta);
    }
    // columns d/dxi, d/deta, d/dzeta of (x, y, z)
    void jac(double eta, double zeta, double M[3][3]) const
    {
        const double yy = y(eta), ye = y_eta(eta);
        const double sz = std::sin(2 * PI * zeta / Lz), cz = std::cos(2 * PI * zeta / Lz);
        M[0][0] = 1;
        M[0][1] = a * 2 * PI * std::cos(2 * PI * yy) * ye * sz;
        M[0][2] = a * std::sin(2 * PI * yy) * (2 * PI / Lz) * cz;
        M[1][0] = 0; M[1][1] = ye; M[1][2] = 0;
        M[2][0] = 0; M[2][1] = 0; M[2][2] = 1;
    }
};

// metrics d(xi,eta,zeta)/d(x,y,z) = M^-1 by cofactors, and J = 1/det M
void invert(const double M[3][3], double m[9], double &J)
{
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    m[0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    m[1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    m[2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    m[3] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    m[4] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    m[5] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    m[6] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    m[7] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    m[8] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    J = 1. / det;
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
double unit(uint64_t &s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

}  // namespace

extern "C" {

void cgli_channel_defaults(cgli_channel *ch, double re_tau)
{
    ch->Lx = 4 * PI;
    ch->Lz = 2 * PI;
    ch->beta = 2.0;
    ch->skew = 0.0;
    ch->mach = 0.5;
    ch->re_tau = re_tau;
    // centreline Reynolds number from U_c/u_tau of turbulent channel data
    ch->re_c = re_tau <= 200 ? 3300.0 : re_tau <= 450 ? 7900.0 : 12500.0;
    ch->Pr = 0.72;
    ch->suth_S = 110.4 / 288.15;
    ch->cfl = 0.5;
    ch->noise = 0.1;
    // preconditioning floor mach^2: with the reference's 1e-4 the inner
    // iterations diverge next to the walls, where the seeded noise leaves
    // cells with |u| ~ 0.02 and beta pinned at the floor
    ch->precond_e = 0.0;
    ch->seed = 20261203ULL;
    ch->xi_mode = 1;
}

int cgli_channel_config(const cgli_channel *ch, int nx, int ny, int nz, cgli_config *c)
{
    if (!ch || !c || nx < 4 || ny < 6 || nz < 4) return CGLI_E_ARG;
    cgli_config_reference(c, nx, ny, nz);
    c->first = 0; c->last = 0; c->x_periodic = 1;
    c->gstart = 0; c->gend0 = nx - 1; c->nx_global = nx;
    c->sponge = 0;
    c->xi_mode = ch->xi_mode;
    c->dxi = ch->Lx / nx; c->deta = 1.0 / ny; c->dzeta = ch->Lz / nz;
    c->high = 2.0;
    c->K = 1.4;
    c->R = 1.0 / (c->K * ch->mach * ch->mach);   // rho = T = 1, U_c = 1
    c->Cv = c->R / (c->K - 1);
    c->mu_L = 1.0 / ch->re_c;
    c->Pr_L = ch->Pr;
    c->suth_T = 1.0;
    c->suth_S = ch->suth_S;
    c->rho_ref = 1.0;
    c->grav = 0.0;
    const double utau = ch->re_tau / ch->re_c;   // rho = h = 1
    c->f_const = utau * utau;                    // driving pressure gradient
    c->f_dplusgs = c->f_const;
    c->E_ref = 0.0;
    return CGLI_OK;
}

int cgli_channel_init(cgli_ctx *ctx, const cgli_channel *ch, double *deltaT)
{
    if (!ctx || !ch) return CGLI_E_ARG;
    const cgli_config *cfg = ctx_config(ctx);
    const hipStream_t s = ctx_stream(ctx);
    const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    const long Ym = ny + 4, Zm = nz + 4, plane = Ym * Zm, N = (long)(nx + 6) * plane;
    const int NY = ny + 1, NZ = nz + 1, gend = nx + 2;
    const Mapping map{ch->Lz, ch->beta, ch->skew};
    // metric planes: cell, u, v, w sets; eta/zeta offsets per set
    const double eoff[4] = {1.5, 0.5, 1.0, 0.5}, zoff[4] = {1.5, 0.5, 0.5, 1.0};
    const int jid[4] = {CGLI_J, CGLI_J_U, CGLI_J_V, CGLI_J_W};
    const int mid[4] = {CGLI_MET, CGLI_MET_U, CGLI_MET_V, CGLI_MET_W};
    std::vector<double> pl(10 * plane);
    std::vector<double> cellm(10 * plane);
    for (int set = 0; set < 4; set++) {
        for (int j = 0; j < Ym; j++)
            for (int k = 0; k < Zm; k++) {
                double M[3][3], m[9], J;
                map.jac((j - eoff[set]) * cfg->deta, (k - zoff[set]) * cfg->dzeta, M);
                invert(M, m, J);
                pl[j * Zm + k] = J;
                for (int q = 0; q < 9; q++) pl[(1 + q) * plane + j * Zm + k] = m[q];
            }
        if (set == 0) cellm = pl;
        for (int q = 0; q < 10; q++) {
            const int id = q == 0 ? jid[set] : mid[set] + q - 1;
            double *d = nullptr;
            int r = ctx_field(ctx, id, &d);
            if (r) return r;
            if (hipMemcpyAsync(d, pl.data() + q * plane, sizeof(double) * plane, hipMemcpyHostToDevice, s) !=
                hipSuccess)
                return CGLI_E_HIP;
            launch_replicate_planes(d, plane, nx + 6, s);
            if (hipStreamSynchronize(s) != hipSuccess) return CGLI_E_HIP;
            ctx_mark(ctx, id, 1);
        }
    }
    // y coordinate of the cell centres (Filter input).  Ghost rows mirror the
    // first and last interior rows about the walls y = -1 and y = 1, the
    // rule of Grid.cpp:88-89 (whose walls are y = 0 and y = high).
    {
        for (int j = 0; j < Ym; j++) {
            double yy;
            if (j < 2) yy = -2.0 - map.y((3 - j - 1.5) * cfg->deta);
            else if (j > NY) yy = 2.0 - map.y((2 * NY + 1 - j - 1.5) * cfg->deta);
            else yy = map.y((j - 1.5) * cfg->deta);
            for (int k = 0; k < Zm; k++) pl[j * Zm + k] = yy;
        }
        double *d = nullptr;
        int r = ctx_field(ctx, CGLI_YPOINT, &d);
        if (r) return r;
        if (hipMemcpyAsync(d, pl.data(), sizeof(double) * plane, hipMemcpyHostToDevice, s) != hipSuccess)
            return CGLI_E_HIP;
        launch_replicate_planes(d, plane, nx + 6, s);
        ctx_mark(ctx, CGLI_YPOINT, 1);
    }
    // initial state: laminar u = 1 - y^2 plus uniform noise on u, v, w; rho =
    // T = 1, p = R.  One splitmix64 stream per velocity component, drawn over
    // the interior cells with i fastest, then j, then k.
    std::vector<double> q5(5 * N, 0.0);
    uint64_t st[3] = {ch->seed * 3 + 0, ch->seed * 3 + 1, ch->seed * 3 + 2};
    std::vector<double> nzs(3 * (size_t)nx * ny * nz);
    const size_t ncell = (size_t)nx * ny * nz;
    for (int q = 0; q < 3; q++)
        for (size_t t = 0; t < ncell; t++) nzs[q * ncell + t] = ch->noise * (2 * unit(st[q]) - 1);
    const double K = cfg->K, p0 = cfg->R;
    double lam_max = 0.0;
    for (int k = 2; k <= NZ; k++)
        for (int j = 2; j <= NY; j++) {
            const double yy = map.y((j - 1.5) * cfg->deta);
            const long pj = (long)j * Zm + k;
            const double J = cellm[pj];
            const double *m = nullptr;
            double mm[9];
            for (int q = 0; q < 9; q++) mm[q] = cellm[(1 + q) * plane + pj];
            m = mm;
            for (int i = 3; i <= gend; i++) {
                const size_t t = (size_t)(i - 3) + (size_t)nx * ((j - 2) + (size_t)ny * (k - 2));
                const double u = (1 - yy * yy) + nzs[t], v = nzs[ncell + t], w = nzs[2 * ncell + t];
                const double rho = 1.0;
                const long c = i * plane + pj;
                q5[c] = rho / J;
                q5[N + c] = rho * u / J;
                q5[2 * N + c] = rho * v / J;
                q5[3 * N + c] = rho * w / J;
                q5[4 * N + c] = (p0 / (K - 1) + 0.5 * rho * (u * u + v * v + w * w)) / J;
                const double cs = std::sqrt(K * p0 / rho);
                const double U = m[0] * u + m[1] * v + m[2] * w;
                const double V = m[3] * u + m[4] * v + m[5] * w;
                const double W = m[6] * u + m[7] * v + m[8] * w;
                const double lam = (std::fabs(U) + cs * std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])) / cfg->dxi +
                                   (std::fabs(V) + cs * std::sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5])) / cfg->deta +
                                   (std::fabs(W) + cs * std::sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8])) / cfg->dzeta;
                lam_max = std::max(lam_max, lam);
            }
        }
    for (int lvl = 0; lvl < 3; lvl++)
        for (int q = 0; q < 5; q++) {
            const int id = (lvl == 0 ? CGLI_U_ : lvl == 1 ? CGLI_U : CGLI_UQ) + q;
            double *d = nullptr;
            int r = ctx_field(ctx, id, &d);
            if (r) return r;
            if (hipMemcpyAsync(d, q5.data() + q * N, sizeof(double) * N, hipMemcpyHostToDevice, s) != hipSuccess)
                return CGLI_E_HIP;
            if (hipStreamSynchronize(s) != hipSuccess) return CGLI_E_HIP;
            ctx_mark(ctx, id, 1);
        }
    ctx_channel(ctx).e = ch->precond_e > 0 ? ch->precond_e : ch->mach * ch->mach;
    if (deltaT) *deltaT = ch->cfl / lam_max;
    return CGLI_OK;
}

int cgli_channel_ghosts(cgli_ctx *ctx)
{
    // isothermal no-slip walls at T = 1 (p R T with R = p0 here: T_wall = 1
    // in units of the reference temperature), then periodic z, then x
    const cgli_config *cfg = ctx_config(ctx);
    int r = cgli_bc_y(ctx, 1.0, 1.0, -2147483647 - 1, 2147483647);
    if (r) return r;
    (void)cfg;
    r = cgli_bc_z(ctx);
    if (r) return r;
    return cgli_bc_x_periodic(ctx);
}

}  // extern "C"

namespace {

// the start of an inner iteration: ghost fill and x fluxes
int queue_prefix(cgli_ctx *ctx)
{
    const double e = ctx_channel(ctx).e;
    int r = cgli_channel_ghosts(ctx);
    if (!r) r = cgli_flux_x(ctx, 1.0, e);
    return r;
}

// one inner iteration queued on the context stream: ghosts, fluxes and
// viscous terms (unless the right-hand side is already in place; prefix:
// the ghost fill and x fluxes are already queued), then the dual-time
// update up to the norm's copy to the host.  later: not the step's first
// iteration, so U_'s interior is the previous iteration's update.
int queue_iteration(cgli_ctx *ctx, double deltaT, int sweeps, bool reuse, bool later, bool prefix = false)
{
    const double e = ctx_channel(ctx).e;
    int r = CGLI_OK;
    if (!reuse) {
        if (!prefix) r = queue_prefix(ctx);
        if (!r) r = cgli_flux_y(ctx, 1.0, e);
        if (!r) r = cgli_flux_z(ctx, 1.0, e);
        if (!r) r = cgli_viscous(ctx);
    }
    if (!r) r = ctx_dts_async(ctx, deltaT, e, sweeps, 0, reuse, later);
    return r;
}

// queue one inner iteration, from the graph of its key when graphs are on:
// a key's first occurrence runs eagerly (it allocates what the iteration
// needs), its second is captured; a graph gone stale is captured again
int run_iteration(cgli_ctx *ctx, ChannelState &cs, double deltaT, int sweeps, bool reuse, bool later, bool fp32,
                  bool prefix)
{
    if (!cs.graph || reuse || ctx_timing_on(ctx)) return queue_iteration(ctx, deltaT, sweeps, reuse, later, prefix);
    ChannelState::GraphSlot *slot = nullptr;
    for (auto &q : cs.graphs)
        if (q.fp32 == (int)fp32 && q.dt == deltaT && q.sweeps == sweeps && q.later == (int)later) slot = &q;
    if (!slot) {
        cs.graphs.push_back({(int)fp32, deltaT, sweeps, (int)later, nullptr});
        return queue_iteration(ctx, deltaT, sweeps, reuse, later);
    }
    if (slot->g) {
        const int r = cgli_graph_launch(slot->g);
        if (r != CGLI_E_STATE) return r;
        // stale (the context changed): eagerly now, captured next time
        cgli_graph_destroy(slot->g);
        slot->g = nullptr;
        return queue_iteration(ctx, deltaT, sweeps, reuse, later);
    }
    int r = cgli_graph_begin(ctx);
    if (r) return r;
    r = queue_iteration(ctx, deltaT, sweeps, reuse, later);
    cgli_graph *G = nullptr;
    const int re = cgli_graph_end(ctx, &G);
    if (!r) r = re;
    if (r) {
        if (G) cgli_graph_destroy(G);
        // CGLI_E_STATE: the iteration needed an allocation (a new schedule's
        // buffers); nothing ran, so run it eagerly and capture next time
        return r == CGLI_E_STATE ? queue_iteration(ctx, deltaT, sweeps, reuse, later) : r;
    }
    slot->g = G;
    return cgli_graph_launch(G);
}

// residual sums of squares of the current U_ (ghosts, fluxes, viscous
// terms and the right-hand side of the inner iteration)
int guess_residual(cgli_ctx *ctx, double deltaT, double out[5])
{
    const double e = ctx_channel(ctx).e;
    int r = cgli_channel_ghosts(ctx);
    if (!r) r = cgli_flux_x(ctx, 1.0, e);
    if (!r) r = cgli_flux_y(ctx, 1.0, e);
    if (!r) r = cgli_flux_z(ctx, 1.0, e);
    if (!r) r = cgli_viscous(ctx);
    if (!r) r = ctx_rhs_sumsq(ctx, deltaT, e, out);
    return r;
}

// The network's guess is checked against the linear extrapolation (one
// more residual evaluation) on its first use, on every NET_CHECK_EVERY-th
// step and on every step after a rejection; in between it is taken as it is.
constexpr int NET_CHECK_EVERY = 8;

// U_ at the start of a step (after the time shift U_ = U = U^n).  *ready
// is set when the last evaluation left fluxes and right-hand side of the
// chosen guess in place, so that the first inner iteration can reuse them.
int initial_guess(cgli_ctx *ctx, double deltaT, ChannelState &cs, bool *ready)
{
    cs.used_net = 0;
    *ready = false;
    if (cs.guess == 0) return CGLI_OK;
    const Geo &g = ctx_geo(ctx);
    const hipStream_t s = ctx_stream(ctx);
    double *u_[5], *u[5], *uq[5];
    for (int q = 0; q < 5; q++)
        if (ctx_field(ctx, CGLI_U_ + q, &u_[q]) || ctx_field(ctx, CGLI_U + q, &u[q]) ||
            ctx_field(ctx, CGLI_UQ + q, &uq[q]))
            return CGLI_E_NOMEM;
    double *J = nullptr;
    if (ctx_field(ctx, CGLI_J, &J)) return CGLI_E_NOMEM;
    const F5 U_{{u_[0], u_[1], u_[2], u_[3], u_[4]}};
    const C5 U{{u[0], u[1], u[2], u[3], u[4]}}, Uq{{uq[0], uq[1], uq[2], uq[3], uq[4]}};
    const float *prev = cs.guess >= 2 ? net_prev_record(cs.net, ctx) : nullptr;
    if (cs.guess == 2 && prev && net_ready(cs.net, ctx)) {
        const float *Y = nullptr, *rs = nullptr;
        int r = CGLI_OK;
        if (!net_take_ahead(cs.net, ctx, &Y, &rs)) r = net_predict(cs.net, s, &Y, &rs);
        if (r) return r;
        if (cs.verify || cs.since_check + 1 >= NET_CHECK_EVERY) {
            // the extrapolated guess first, the network's last: when the
            // network's is taken its right-hand side is still in place
            net_guess(g, U, Uq, J, nullptr, nullptr, nullptr, U_, s);
            r = guess_residual(ctx, deltaT, cs.r_ext);
            if (r) return r;
            net_guess(g, U, Uq, J, prev, Y, rs, U_, s);
            r = guess_residual(ctx, deltaT, cs.r_net);
            if (r) return r;
            bool worse = false;
            for (int q = 0; q < 5; q++) worse = worse || cs.r_net[q] > cs.r_ext[q];
            if (worse) {
                net_guess(g, U, Uq, J, nullptr, nullptr, nullptr, U_, s);
            } else {
                cs.used_net = 1;
                *ready = true;
            }
            cs.verify = worse;
            cs.since_check = 0;
        } else {
            net_guess(g, U, Uq, J, prev, Y, rs, U_, s);
            cs.used_net = 1;
            cs.since_check++;
        }
    } else if (cs.guess == 3 && prev) {
        net_guess(g, U, Uq, J, prev, nullptr, nullptr, U_, s);
    } else {
        net_guess(g, U, Uq, J, nullptr, nullptr, nullptr, U_, s);
    }
    return hipGetLastError() == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

}  // namespace

extern "C" {

int cgli_channel_set_rule(cgli_ctx *ctx, int rule)
{
    if (!ctx || rule < 0 || rule > 1) return CGLI_E_ARG;
    ChannelState &cs = ctx_channel(ctx);
    cs.rule = rule;
    for (int v = 0; v < 5; v++) cs.E[v] = 0.0;
    return CGLI_OK;
}

int cgli_channel_rule_state(cgli_ctx *ctx, double E[5], int set)
{
    if (!ctx || !E) return CGLI_E_ARG;
    ChannelState &cs = ctx_channel(ctx);
    for (int v = 0; v < 5; v++) {
        if (set) cs.E[v] = E[v];
        else E[v] = cs.E[v];
    }
    return CGLI_OK;
}

int cgli_channel_set_graph(cgli_ctx *ctx, int on)
{
    if (!ctx) return CGLI_E_ARG;
    ChannelState &cs = ctx_channel(ctx);
    cs.graph = on != 0;
    if (!cs.graph) channel_drop_graphs(cs);
    return CGLI_OK;
}

int cgli_channel_set_guess(cgli_ctx *ctx, int guess)
{
    if (!ctx || guess < 0 || guess > 3) return CGLI_E_ARG;
    ChannelState &cs = ctx_channel(ctx);
    cs.guess = guess;
    cs.verify = true;   // the network's guess is checked on its next use
    cs.since_check = 0;
    return CGLI_OK;
}

int cgli_channel_attach_net(cgli_ctx *ctx, cgli_net *net)
{
    if (!ctx) return CGLI_E_ARG;
    ChannelState &cs = ctx_channel(ctx);
    cs.net = net;
    cs.verify = true;
    cs.since_check = 0;
    return CGLI_OK;
}

int cgli_channel_guess_info(cgli_ctx *ctx, int *used_net, double r_net[5], double r_ext[5])
{
    if (!ctx) return CGLI_E_ARG;
    const ChannelState &cs = ctx_channel(ctx);
    if (used_net) *used_net = cs.used_net;
    for (int v = 0; v < 5; v++) {
        if (r_net) r_net[v] = cs.r_net[v];
        if (r_ext) r_ext[v] = cs.r_ext[v];
    }
    return CGLI_OK;
}

int cgli_channel_step(cgli_ctx *ctx, double deltaT, double tol, int max_inner, int sweeps, int *inner,
                      double er_first[5], double er_last[5])
{
    if (!ctx || max_inner < 1) return CGLI_E_ARG;
    const cgli_config *cfg = ctx_config(ctx);
    const hipStream_t s = ctx_stream(ctx);
    ChannelState &cs = ctx_channel(ctx);
    double e1[5] = {0, 0, 0, 0, 0}, er[5] = {0, 0, 0, 0, 0};
    int it = 0;
    bool ready = false;
    {
        ctx_set_fp32(ctx, cfg->precision == 1);
        const int r = initial_guess(ctx, deltaT, cs, &ready);
        if (r) return r;
    }
    // Mixed precision (cfg->precision == 1): FP32 residual kernels while the
    // drop is above 100 x tol and still falling, FP64 afterwards.  The
    // residual's rounding bounds the drop an iteration can reach (about
    // 2e-5 with FP32 stencils on the channel), so the last iterations run
    // in FP64 and the converged state is the FP64 fixed point.
    bool fp32 = cfg->precision == 1;
    int n32 = 0;
    double prev = 1.0;
    // Eagerly launched iterations queue the next iteration's ghost fill and
    // x fluxes before the host reads this iteration's norm, so that the
    // device works while the host decides.  Should the step stop there, that
    // work only refilled the ghosts of the final state and rewrote x fluxes
    // that the next use recomputes.
    bool prefix = false;
    for (it = 1; it <= max_inner; it++) {
        ctx_set_fp32(ctx, fp32);
        if (fp32) n32++;
        // the first iteration reuses the residual evaluation of the guess
        // check when the network's guess was taken
        const bool reuse = ready && it == 1;
        int r = run_iteration(ctx, cs, deltaT, sweeps, reuse, it > 1, fp32, prefix);
        prefix = false;
        if (!r && it < max_inner && !(cs.graph && !ctx_timing_on(ctx))) {
            r = queue_prefix(ctx);
            prefix = r == CGLI_OK;
        }
        if (!r) r = ctx_dts_wait(ctx, er);
        if (r) return r;
        if (it == 1)
            for (int v = 0; v < 5; v++) e1[v] = er[v];
        // rule 1 normalises by the running maximum (main.cpp:287-304)
        const double *ref = e1;
        if (cs.rule == 1) {
            for (int v = 0; v < 5; v++) cs.E[v] = std::max(cs.E[v], er[v]);
            ref = cs.E;
        }
        double drop = 0.0;
        for (int v = 0; v < 5; v++) drop = std::max(drop, ref[v] > 0 ? er[v] / ref[v] : 0.0);
        if (drop <= tol) break;
        if (fp32 && (drop <= 100 * tol || (it > 2 && drop > 0.85 * prev))) {
            fp32 = false;
            prefix = false;   // the queued x fluxes were FP32
        }
        prev = drop;
    }
    // every step ends with the ghosts of its final state filled (a queued
    // prefix has done it), whichever way its iterations were launched
    if (!prefix) {
        const int r = cgli_channel_ghosts(ctx);
        if (r) return r;
    }
    ctx_set_fp32(ctx, cfg->precision == 1);
    ctx_set_last_fp32(ctx, n32);
    if (it > max_inner) it = max_inner;
    // time-level shift Uq <- U, U <- U_ (main.cpp:364-386)
    const long N = (long)(cfg->nx + 6) * (cfg->ny + 4) * (cfg->nz + 4);
    for (int q = 0; q < 5; q++) {
        double *u_ = nullptr, *u = nullptr, *uq = nullptr;
        ctx_field(ctx, CGLI_U_ + q, &u_);
        ctx_field(ctx, CGLI_U + q, &u);
        ctx_field(ctx, CGLI_UQ + q, &uq);
        launch_copy_field(uq, u, N, s);
        launch_copy_field(u, u_, N, s);
    }
    if (cs.net) {
        int r = net_record(cs.net, ctx);
        // the next step's prediction runs on the network's stream from here
        if (!r && cs.guess == 2) r = net_predict_ahead(cs.net, ctx);
        if (r) return r;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return CGLI_E_HIP;
    if (inner) *inner = it;
    if (er_first)
        for (int v = 0; v < 5; v++) er_first[v] = e1[v];
    if (er_last)
        for (int v = 0; v < 5; v++) er_last[v] = er[v];
    return CGLI_OK;
}

}  // extern "C"
