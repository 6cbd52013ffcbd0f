// learned.cpp -- the host side of the learned initial guess: parameters,
// Adam state, the snapshot ring, training and whole-grid prediction, over
// the kernels of cnn.hip.  No counterpart in the reference, whose steps
// start from U_ = U^n (main.cpp:378-382).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "cgli_internal.hpp"

using namespace cgli;

namespace {

struct Layer { int cin, cin_pad, cout, cout_pad; long w, b; };

// parameter layout: W[Cout_pad][27][Cin_pad] then b[Cout_pad] per layer
constexpr Layer LAYERS[4] = {
    {15, 16, 32, 32, 0, 13824},
    {32, 32, 32, 32, 13856, 41504},
    {32, 32, 32, 32, 41536, 69184},
    {32, 32, 5, 16, 69216, 83040},
};
constexpr long NPARAM = 83056;
static_assert(NPARAM == CGLI_NET_NPARAM, "parameter count");

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
double unit(uint64_t &s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

template <class T>
int grow(T *&p, long &cap, long n)
{
    if (cap >= n) return CGLI_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * n) != hipSuccess) return CGLI_E_NOMEM;
    p = static_cast<T *>(q);
    cap = n;
    return CGLI_OK;
}

// activation and gradient buffers for nv voxels
struct Work {
    long nv = 0;
    uint16_t *X0 = nullptr; long cX0 = 0;     // [nv][16] BF16
    uint16_t *A[3] = {}; long cA[3] = {};     // [nv][32] BF16
    float *Z[3] = {}; long cZ[3] = {};        // [nv][32] FP32 pre-activations
    float *Y = nullptr; long cY = 0;          // [nv][16]
    float *T = nullptr; long cT = 0;          // [nv][5]
    unsigned char *mask = nullptr; long cM = 0;
    float *dY = nullptr; long cdY = 0;        // [nv][16]
    uint16_t *dZb = nullptr; long cdZb = 0;   // [nv][32] BF16
    float *dA = nullptr; long cdA = 0;        // [nv][32]
    float *dZf = nullptr; long cdZf = 0;      // [nv][32]
    float *part = nullptr; long cpart = 0;    // weight-gradient partials
    double *lpart = nullptr; long clpart = 0; // loss partials
    double *bpart = nullptr; long cbpart = 0; // bias-gradient partials
    float *xin = nullptr; long cxin = 0;      // host input staging [nv][15]
};

int wgrad_blocks(long nv) { return (int)std::max(1L, std::min(256L, nv / 2048)); }
int mse_blocks(long nv) { return (int)std::max(1L, std::min(512L, (nv * 16 + 2047) / 2048)); }

}  // namespace

struct EventPair { hipEvent_t a = nullptr, b = nullptr; };

struct cgli_net {
    int device = 0;
    hipStream_t stream = nullptr;
    float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *g = nullptr;
    uint16_t *wb = nullptr;       // BF16 copy of w
    uint16_t *wt[4] = {};         // transposed BF16 weights of layers 1..3
    long adam_t = 0;
    bool trained = false;
    // snapshot ring and each record's inverse scales (1/s1, 1/sd, 1/s2)
    float *ring = nullptr;
    float *rscale = nullptr;
    int cap = 0;
    long nrec = 0, vol = 0;
    int nx = 0, ny = 0, nz = 0;
    int *org = nullptr;           // device crop origins
    long corg = 0;
    double *loss_dev = nullptr;   // per-step losses
    long closs = 0;
    double *part10 = nullptr;    // partial sums of the record scales
    Work wk;
    hipEvent_t ev[4] = {};
    // a prediction started at the end of a step on this stream, for the
    // next step's guess: valid while ahead_rec equals nrec
    hipEvent_t ahead_in = nullptr, ahead_out = nullptr;
    long ahead_rec = -1;
    const float *ahead_Y = nullptr, *ahead_rs = nullptr;
    std::vector<EventPair> pred_ev;   // one pair per prediction since the last timing read
    size_t pred_used = 0;
    double predict_ms = 0, train_ms = 0;
    long n_predict = 0, n_train = 0;
};

namespace {

EventPair &timing_pair(std::vector<EventPair> &pool, size_t &used)
{
    if (used == pool.size()) {
        EventPair e;
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
        pool.push_back(e);
    }
    return pool[used++];
}

// fold the recorded prediction times into the totals
void collect_predict_times(cgli_net *n)
{
    for (size_t q = 0; q < n->pred_used; q++) {
        float ms = 0;
        if (hipEventSynchronize(n->pred_ev[q].b) == hipSuccess &&
            hipEventElapsedTime(&ms, n->pred_ev[q].a, n->pred_ev[q].b) == hipSuccess)
            n->predict_ms += ms;
        n->n_predict++;
    }
    n->pred_used = 0;
}

// BF16 copy and transposed weights from the FP32 master
void refresh_bf16(cgli_net *n, hipStream_t s)
{
    net_to_bf16(n->w, n->wb, NPARAM, s);
    for (int l = 1; l < 4; l++)
        net_transpose_w(n->wb + LAYERS[l].w, n->wt[l], LAYERS[l].cin_pad, LAYERS[l].cout_pad, s);
}

void transpose_all(cgli_net *n, hipStream_t s)
{
    for (int l = 1; l < 4; l++)
        net_transpose_w(n->wb + LAYERS[l].w, n->wt[l], LAYERS[l].cin_pad, LAYERS[l].cout_pad, s);
}

int ensure_work(cgli_net *n, long nv, bool train)
{
    Work &k = n->wk;
    int r = grow(k.X0, k.cX0, nv * 16);
    // inference ping-pongs A[0] and A[1]; training keeps all three
    for (int l = 0; l < (train ? 3 : 2) && !r; l++) r = grow(k.A[l], k.cA[l], nv * 32);
    if (!r) r = grow(k.Y, k.cY, nv * 16);
    if (train) {
        for (int l = 0; l < 3 && !r; l++) r = grow(k.Z[l], k.cZ[l], nv * 32);
        if (!r) r = grow(k.T, k.cT, nv * 5);
        if (!r) r = grow(k.mask, k.cM, nv);
        if (!r) r = grow(k.dY, k.cdY, nv * 16);
        if (!r) r = grow(k.dZb, k.cdZb, nv * 32);
        if (!r) r = grow(k.dA, k.cdA, nv * 32);
        if (!r) r = grow(k.dZf, k.cdZf, nv * 32);
        if (!r) r = grow(k.part, k.cpart, (long)wgrad_blocks(nv) * 27 * 32 * 32);
        if (!r) r = grow(k.lpart, k.clpart, (long)mse_blocks(nv));
        if (!r) r = grow(k.bpart, k.cbpart, (long)NET_BIAS_BLOCKS * 32);
    }
    return r;
}

// forward pass over X0 (already packed); keeps Z/A for training when train
int forward(cgli_net *n, const NetDims &m, int per, bool train, hipStream_t s)
{
    Work &k = n->wk;
    const void *in = k.X0;
    for (int l = 0; l < 3; l++) {
        const Layer &L = LAYERS[l];
        // inference ping-pongs A[0] and A[1]
        void *out = train ? k.A[l] : k.A[l & 1];
        if (net_conv(L.cin_pad, L.cout_pad, true, in, n->wb + L.w, n->w + L.b, train ? k.Z[l] : nullptr, out, m,
                     L.cout, per, s))
            return CGLI_E_ARG;
        in = out;
    }
    const Layer &L = LAYERS[3];
    if (net_conv(L.cin_pad, L.cout_pad, false, in, n->wb + L.w, n->w + L.b, k.Y, nullptr, m, L.cout, per, s))
        return CGLI_E_ARG;
    return CGLI_OK;
}

// masked MSE and the full backward pass into n->g; loss into *loss_dev
int backward(cgli_net *n, const NetDims &m, float inv, double *loss_dev, hipStream_t s)
{
    Work &k = n->wk;
    const long nv = (long)m.B * m.D * m.H * m.W;
    const int mb = mse_blocks(nv), wb = wgrad_blocks(nv);
    net_mse(k.Y, k.T, k.mask, nv, inv, k.dY, k.lpart, mb, loss_dev, inv, s);
    // layer 4 (no activation): dz = dY
    net_to_bf16(k.dY, k.dZb, nv * 16, s);
    for (int l = 3; l >= 0; l--) {
        const Layer &L = LAYERS[l];
        const void *xin = l == 0 ? (const void *)k.X0 : (const void *)k.A[l - 1];
        if (net_wgrad(L.cin_pad, L.cout_pad, k.dZb, xin, m, wb, k.part, L.cout, L.cin, n->g + L.w, s))
            return CGLI_E_ARG;
        net_bias_grad(l == 3 ? k.dY : k.dZf, nv, L.cout_pad, L.cout, n->g + L.b, k.bpart, s);
        if (l == 0) break;
        // dA_{l-1} = conv(dz_l, W_l^T), then dz_{l-1} = dA * GELU'(z_{l-1})
        if (net_conv(L.cout_pad, L.cin_pad, false, k.dZb, n->wt[l], nullptr, k.dA, nullptr, m, L.cin, 0, s))
            return CGLI_E_ARG;
        net_gelu_back(k.dA, k.Z[l - 1], k.dZb, k.dZf, nv * 32, s);
    }
    return CGLI_OK;
}

void adam_step(cgli_net *n, float lr, hipStream_t s)
{
    const double b1 = 0.9, b2 = 0.999;
    n->adam_t++;
    const float c1 = (float)(1.0 - std::pow(b1, (double)n->adam_t));
    const float c2 = (float)(1.0 - std::pow(b2, (double)n->adam_t));
    net_adam(n->w, n->m1, n->m2, n->g, n->wb, NPARAM, lr, (float)b1, (float)b2, 1e-8f, c1, c2, s);
    transpose_all(n, s);
    n->trained = true;
}

// voxels counted by the loss of a crop at j-origin oj (the k_gather rule)
long counted(int oj, int cd, int ch, int cw, int ny, int halo)
{
    const long nd = std::max(0, cd - 2 * halo), nw = std::max(0, cw - 2 * halo);
    int h0 = oj == 0 ? 0 : halo, h1 = oj + ch == ny ? ch : ch - halo;
    return nd * nw * std::max(0, h1 - h0);
}

// inverse scales of the record just stored in slot t % cap
int record_scales(cgli_net *n, long t, hipStream_t s)
{
    const int nblk = 256;
    if (!n->part10 && hipMalloc(&n->part10, sizeof(double) * NET_NSCALE * nblk) != hipSuccess) return CGLI_E_NOMEM;
    const long lo = std::max(0L, t - n->cap + 1);   // oldest record still held with t
    auto slot = [&](long r) { return r >= lo ? n->ring + (r % n->cap) * n->vol * 5 : nullptr; };
    const float *b = slot(t - 1), *c = b ? slot(t - 2) : nullptr;
    net_record_scales(slot(t), b, c, n->vol, n->part10, nblk, n->rscale + (t % n->cap) * NET_NSCALE, s);
    return hipGetLastError() == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

}  // namespace

namespace cgli {

int net_record(cgli_net *net, cgli_ctx *c)
{
    const Geo &g = ctx_geo(c);
    if (!net->ring || g.nx != net->nx || g.ny != net->ny || g.nz != net->nz) return CGLI_E_STATE;
    // a prediction still running reads the ring: the new record waits
    if (net->ahead_rec >= 0 && hipStreamWaitEvent(ctx_stream(c), net->ahead_out, 0) != hipSuccess)
        return CGLI_E_HIP;
    double *u[5], *uq[5], *J = nullptr;
    for (int q = 0; q < 5; q++) {
        if (ctx_field(c, CGLI_U + q, &u[q]) || ctx_field(c, CGLI_UQ + q, &uq[q])) return CGLI_E_NOMEM;
    }
    if (ctx_field(c, CGLI_J, &J)) return CGLI_E_NOMEM;
    net_store(g, C5{{u[0], u[1], u[2], u[3], u[4]}}, C5{{uq[0], uq[1], uq[2], uq[3], uq[4]}}, J,
              net->ring + (net->nrec % net->cap) * net->vol * 5, ctx_stream(c));
    const int r = record_scales(net, net->nrec, ctx_stream(c));
    net->nrec++;
    return r;
}

const float *net_prev_record(cgli_net *net, cgli_ctx *c)
{
    const Geo &g = ctx_geo(c);
    if (!net || !net->ring || g.nx != net->nx || g.ny != net->ny || g.nz != net->nz || net->nrec < 2 ||
        net->cap < 2)
        return nullptr;
    return net->ring + ((net->nrec - 2) % net->cap) * net->vol * 5;
}

int net_ready(cgli_net *net, cgli_ctx *c)
{
    const Geo &g = ctx_geo(c);
    return net && net->trained && net->ring && g.nx == net->nx && g.ny == net->ny &&
           g.nz == net->nz && net->nrec >= 3 && net->cap >= 3 && net->nz % 16 == 0;
}

int net_predict(cgli_net *net, hipStream_t s, const float **Y, const float **rs)
{
    const NetDims m{1, net->nx, net->ny, net->nz};
    int r = ensure_work(net, net->vol, false);
    if (r) return r;
    if (net->pred_used >= 256) collect_predict_times(net);
    EventPair &e = timing_pair(net->pred_ev, net->pred_used);
    hipEventRecord(e.a, s);
    net_gather(net->ring, net->rscale, net->cap, net->vol, net->nx, net->ny, net->nz, nullptr, (int)(net->nrec - 1),
               m, 0, net->wk.X0, nullptr, nullptr, s);
    r = forward(net, m, 1, false, s);
    hipEventRecord(e.b, s);
    if (r) return r;
    *Y = net->wk.Y;
    *rs = net->rscale + ((net->nrec - 1) % net->cap) * NET_NSCALE;
    return hipGetLastError() == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

int net_predict_ahead(cgli_net *net, cgli_ctx *c)
{
    net->ahead_rec = -1;
    if (!net_ready(net, c)) return CGLI_OK;
    if (hipEventRecord(net->ahead_in, ctx_stream(c)) != hipSuccess ||
        hipStreamWaitEvent(net->stream, net->ahead_in, 0) != hipSuccess)
        return CGLI_E_HIP;
    const int r = net_predict(net, net->stream, &net->ahead_Y, &net->ahead_rs);
    if (r) return r;
    if (hipEventRecord(net->ahead_out, net->stream) != hipSuccess) return CGLI_E_HIP;
    net->ahead_rec = net->nrec;
    return CGLI_OK;
}

bool net_take_ahead(cgli_net *net, cgli_ctx *c, const float **Y, const float **rs)
{
    if (!net || net->ahead_rec < 0 || net->ahead_rec != net->nrec) return false;
    if (hipStreamWaitEvent(ctx_stream(c), net->ahead_out, 0) != hipSuccess) return false;
    *Y = net->ahead_Y;
    *rs = net->ahead_rs;
    net->ahead_rec = -1;
    return true;
}

}  // namespace cgli

extern "C" {

int cgli_net_create(int device, unsigned long long seed, cgli_net **out)
{
    if (!out) return CGLI_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CGLI_E_NODEV;
    if (device < 0 || device >= ndev) return CGLI_E_ARG;
    cgli_net *n = new (std::nothrow) cgli_net();
    if (!n) return CGLI_E_NOMEM;
    n->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return CGLI_E_HIP;
    }
    bool ok = true;
    for (float **p : {&n->w, &n->m1, &n->m2, &n->g}) ok = ok && hipMalloc(p, sizeof(float) * NPARAM) == hipSuccess;
    ok = ok && hipMalloc(&n->wb, sizeof(uint16_t) * NPARAM) == hipSuccess;
    for (int l = 1; l < 4 && ok; l++) ok = hipMalloc(&n->wt[l], sizeof(uint16_t) * 27 * 32 * 32) == hipSuccess;
    for (int e = 0; e < 4 && ok; e++) ok = hipEventCreate(&n->ev[e]) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&n->ahead_in, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&n->ahead_out, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        cgli_net_destroy(n);
        return CGLI_E_NOMEM;
    }
    // uniform initialisation: hidden layers +-sqrt(6 / fan_in) (He, for the
    // GELU gain), output layer +-0.1 sqrt(3 / fan_in) so that the first
    // guesses stay near the extrapolation; biases 0
    std::vector<float> h(NPARAM, 0.0f);
    uint64_t st = seed;
    for (int l = 0; l < 4; l++) {
        const Layer &L = LAYERS[l];
        const double fan = 27.0 * L.cin;
        const double bound = l < 3 ? std::sqrt(6.0 / fan) : 0.1 * std::sqrt(3.0 / fan);
        for (int co = 0; co < L.cout; co++)
            for (int tap = 0; tap < 27; tap++)
                for (int ci = 0; ci < L.cin; ci++)
                    h[L.w + ((long)co * 27 + tap) * L.cin_pad + ci] = (float)(bound * (2 * unit(st) - 1));
    }
    const hipStream_t s = n->stream;
    ok = hipMemcpy(n->w, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemsetAsync(n->m1, 0, sizeof(float) * NPARAM, s) == hipSuccess &&
         hipMemsetAsync(n->m2, 0, sizeof(float) * NPARAM, s) == hipSuccess &&
         hipMemsetAsync(n->g, 0, sizeof(float) * NPARAM, s) == hipSuccess;
    if (ok) refresh_bf16(n, s);
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
        cgli_net_destroy(n);
        return CGLI_E_HIP;
    }
    *out = n;
    return CGLI_OK;
}

int cgli_net_destroy(cgli_net *n)
{
    if (!n) return CGLI_E_ARG;
    hipSetDevice(n->device);
    if (n->stream) hipStreamSynchronize(n->stream);
    for (void *p : {(void *)n->w, (void *)n->m1, (void *)n->m2, (void *)n->g, (void *)n->wb, (void *)n->ring,
                    (void *)n->org, (void *)n->loss_dev, (void *)n->rscale, (void *)n->part10})
        if (p) hipFree(p);
    for (int l = 1; l < 4; l++)
        if (n->wt[l]) hipFree(n->wt[l]);
    Work &k = n->wk;
    for (void *p : {(void *)k.X0, (void *)k.A[0], (void *)k.A[1], (void *)k.A[2], (void *)k.Z[0], (void *)k.Z[1],
                    (void *)k.Z[2], (void *)k.Y, (void *)k.T, (void *)k.mask, (void *)k.dY, (void *)k.dZb,
                    (void *)k.dA, (void *)k.dZf, (void *)k.part, (void *)k.lpart, (void *)k.bpart, (void *)k.xin})
        if (p) hipFree(p);
    for (int e = 0; e < 4; e++)
        if (n->ev[e]) hipEventDestroy(n->ev[e]);
    if (n->ahead_in) hipEventDestroy(n->ahead_in);
    if (n->ahead_out) hipEventDestroy(n->ahead_out);
    for (EventPair &e : n->pred_ev) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
    return CGLI_OK;
}

int cgli_net_params(cgli_net *n, float *params, int set)
{
    if (!n || !params) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    const hipStream_t s = n->stream;
    if (hipStreamSynchronize(s) != hipSuccess) return CGLI_E_HIP;
    if (set) {
        if (hipMemcpy(n->w, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess) return CGLI_E_HIP;
        refresh_bf16(n, s);
        n->trained = true;
    } else if (hipMemcpy(params, n->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost) != hipSuccess) {
        return CGLI_E_HIP;
    }
    return hipStreamSynchronize(s) == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

int cgli_net_eval(cgli_net *n, int B, int D, int H, int W, int periodic, const float *x, const float *t,
                  const unsigned char *mask, float *y, double *loss, float *grad)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    if (!n || !x || B < 1 || D < 1 || H < 1 || W < 16 || W % 16) return CGLI_E_ARG;
    if ((loss || grad) && (!t || !mask)) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    const hipStream_t s = n->stream;
    const NetDims m{B, D, H, W};
    const long nv = (long)B * D * H * W;
    const bool train = t && mask;
    int r = ensure_work(n, nv, train);
    if (!r) r = grow(n->wk.xin, n->wk.cxin, nv * 15);
    if (!r && train) r = grow(n->loss_dev, n->closs, 1L);
    if (r) return r;
    Work &k = n->wk;
    if (hipMemcpy(k.xin, x, sizeof(float) * nv * 15, hipMemcpyHostToDevice) != hipSuccess) return CGLI_E_HIP;
    net_pack_input(k.xin, k.X0, nv, s);
    r = forward(n, m, periodic ? 1 : 0, train, s);
    if (r) return r;
    if (train) {
        long cnt = 0;
        for (long v = 0; v < nv; v++) cnt += mask[v] ? 1 : 0;
        if (!cnt) return CGLI_E_ARG;
        if (hipMemcpy(k.T, t, sizeof(float) * nv * 5, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(k.mask, mask, nv, hipMemcpyHostToDevice) != hipSuccess)
            return CGLI_E_HIP;
        r = backward(n, m, (float)(1.0 / (5.0 * cnt)), n->loss_dev, s);
        if (r) return r;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return CGLI_E_HIP;
    if (y) {
        std::vector<float> h(nv * 16);
        if (hipMemcpy(h.data(), k.Y, sizeof(float) * nv * 16, hipMemcpyDeviceToHost) != hipSuccess) return CGLI_E_HIP;
        for (long v = 0; v < nv; v++)
            for (int q = 0; q < 5; q++) y[v * 5 + q] = h[v * 16 + q];
    }
    if (loss && hipMemcpy(loss, n->loss_dev, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return CGLI_E_HIP;
    if (grad && hipMemcpy(grad, n->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost) != hipSuccess)
        return CGLI_E_HIP;
    return hipGetLastError() == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

int cgli_net_adam(cgli_net *n, float lr)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    if (!n || !(lr > 0)) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    adam_step(n, lr, n->stream);
    return hipStreamSynchronize(n->stream) == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

int cgli_net_ring(cgli_net *n, int capacity, int nx, int ny, int nz)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    if (!n || capacity < 4 || nx < 1 || ny < 1 || nz < 16 || nz % 16) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    hipStreamSynchronize(n->stream);
    if (n->ring) hipFree(n->ring);
    if (n->rscale) hipFree(n->rscale);
    n->ring = nullptr;
    n->rscale = nullptr;
    n->cap = 0;
    const long vol = (long)nx * ny * nz;
    if (hipMalloc(&n->ring, sizeof(float) * 5 * vol * capacity) != hipSuccess ||
        hipMalloc(&n->rscale, sizeof(float) * NET_NSCALE * capacity) != hipSuccess)
        return CGLI_E_NOMEM;
    n->cap = capacity;
    n->vol = vol;
    n->nx = nx; n->ny = ny; n->nz = nz;
    n->nrec = 0;
    return CGLI_OK;
}

int cgli_net_push(cgli_net *n, const float *dq)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    if (!n || !dq || !n->ring) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    hipStreamSynchronize(n->stream);
    if (hipMemcpy(n->ring + (n->nrec % n->cap) * n->vol * 5, dq, sizeof(float) * 5 * n->vol,
                  hipMemcpyHostToDevice) != hipSuccess)
        return CGLI_E_HIP;
    const int r = record_scales(n, n->nrec, n->stream);
    n->nrec++;
    if (!r && hipStreamSynchronize(n->stream) != hipSuccess) return CGLI_E_HIP;
    return r;
}

int cgli_net_rewind(cgli_net *n, long nrec)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    // the records nrec-3 .. nrec-1 must still be in the ring
    if (!n || nrec < 0 || nrec > n->nrec || n->nrec - nrec > n->cap - 3) return CGLI_E_ARG;
    n->nrec = nrec;
    return CGLI_OK;
}

int cgli_net_records(cgli_net *n, long *cnt)
{
    if (!n || !cnt) return CGLI_E_ARG;
    *cnt = n->nrec;
    return CGLI_OK;
}

int cgli_net_record_scales(cgli_net *n, long rec, float inv[15])
{
    if (!n || !inv || !n->ring || rec < 0 || rec >= n->nrec || rec < n->nrec - n->cap) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    if (hipStreamSynchronize(n->stream) != hipSuccess ||
        hipMemcpy(inv, n->rscale + (rec % n->cap) * NET_NSCALE, sizeof(float) * NET_NSCALE, hipMemcpyDeviceToHost) !=
            hipSuccess)
        return CGLI_E_HIP;
    return CGLI_OK;
}

int cgli_net_train(cgli_net *n, int steps, int batch, int cd, int ch, int cw, float lr, unsigned long long seed,
                   float *loss)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    const int halo = 4;
    if (!n || steps < 1 || batch < 1 || !(lr > 0) || !n->ring) return CGLI_E_ARG;
    if (cd > n->nx || ch > n->ny || cw > n->nz || cw % 16 || cd <= 2 * halo || cw <= 2 * halo || ch < 1)
        return CGLI_E_ARG;
    const long lo = std::max(0L, n->nrec - n->cap) + 2, hi = n->nrec - 2;   // records t-2 .. t+1 held
    if (hi < lo) return CGLI_E_STATE;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    const hipStream_t s = n->stream;
    const NetDims m{batch, cd, ch, cw};
    const long nv = (long)batch * cd * ch * cw;
    int r = ensure_work(n, nv, true);
    if (!r) r = grow(n->org, n->corg, 4L * batch * steps);
    if (!r) r = grow(n->loss_dev, n->closs, (long)steps);
    if (r) return r;
    std::vector<int> o(4L * batch * steps);
    std::vector<float> inv(steps);
    uint64_t st = seed;
    for (int it = 0; it < steps; it++) {
        long cnt = 0;
        for (int b = 0; b < batch; b++) {
            int *p = &o[4L * (it * batch + b)];
            p[0] = (int)(lo + (long)(unit(st) * (double)(hi - lo + 1)));
            p[0] = std::min<long>(p[0], hi);
            p[1] = (int)(unit(st) * n->nx) % n->nx;
            p[2] = (int)(unit(st) * (n->ny - ch + 1)) % (n->ny - ch + 1);
            p[3] = (int)(unit(st) * n->nz) % n->nz;
            cnt += counted(p[2], cd, ch, cw, n->ny, halo);
        }
        if (!cnt) return CGLI_E_ARG;
        inv[it] = (float)(1.0 / (5.0 * cnt));
    }
    if (hipMemcpy(n->org, o.data(), sizeof(int) * o.size(), hipMemcpyHostToDevice) != hipSuccess) return CGLI_E_HIP;
    hipEventRecord(n->ev[2], s);
    for (int it = 0; it < steps && !r; it++) {
        net_gather(n->ring, n->rscale, n->cap, n->vol, n->nx, n->ny, n->nz, n->org + 4L * it * batch, -1, m, halo,
                   n->wk.X0, n->wk.T, n->wk.mask, s);
        r = forward(n, m, 0, true, s);
        if (!r) r = backward(n, m, inv[it], n->loss_dev + it, s);
        if (!r) adam_step(n, lr, s);
    }
    hipEventRecord(n->ev[3], s);
    if (r) return r;
    if (hipEventSynchronize(n->ev[3]) != hipSuccess) return CGLI_E_HIP;
    float ms = 0;
    hipEventElapsedTime(&ms, n->ev[2], n->ev[3]);
    n->train_ms += ms;
    n->n_train += steps;
    if (loss) {
        std::vector<double> h(steps);
        if (hipMemcpy(h.data(), n->loss_dev, sizeof(double) * steps, hipMemcpyDeviceToHost) != hipSuccess)
            return CGLI_E_HIP;
        for (int it = 0; it < steps; it++) loss[it] = (float)h[it];
    }
    return hipGetLastError() == hipSuccess ? CGLI_OK : CGLI_E_HIP;
}

int cgli_net_predict(cgli_net *n, float *x, float *y)
{
    if (n) n->ahead_rec = -1;   // its work buffers, weights or ring change
    if (!n || !n->ring || n->nrec < 3) return CGLI_E_STATE;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    const float *Y = nullptr, *rs = nullptr;
    int r = net_predict(n, n->stream, &Y, &rs);
    if (r) return r;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return CGLI_E_HIP;
    const long nv = n->vol;
    if (x) {
        std::vector<uint16_t> h(nv * 16);
        if (hipMemcpy(h.data(), n->wk.X0, sizeof(uint16_t) * nv * 16, hipMemcpyDeviceToHost) != hipSuccess)
            return CGLI_E_HIP;
        for (long v = 0; v < nv; v++)
            for (int q = 0; q < 15; q++) {
                const uint32_t u = (uint32_t)h[v * 16 + q] << 16;
                std::memcpy(&x[v * 15 + q], &u, 4);
            }
    }
    if (y) {
        std::vector<float> h(nv * 16);
        if (hipMemcpy(h.data(), Y, sizeof(float) * nv * 16, hipMemcpyDeviceToHost) != hipSuccess) return CGLI_E_HIP;
        for (long v = 0; v < nv; v++)
            for (int q = 0; q < 5; q++) y[v * 5 + q] = h[v * 16 + q];
    }
    return CGLI_OK;
}

int cgli_net_timing(cgli_net *n, double *predict_ms, long *n_predict, double *train_ms, long *n_train)
{
    if (!n) return CGLI_E_ARG;
    if (hipSetDevice(n->device) != hipSuccess) return CGLI_E_HIP;
    collect_predict_times(n);
    if (predict_ms) *predict_ms = n->predict_ms;
    if (n_predict) *n_predict = n->n_predict;
    if (train_ms) *train_ms = n->train_ms;
    if (n_train) *n_train = n->n_train;
    return CGLI_OK;
}

}  // extern "C"
