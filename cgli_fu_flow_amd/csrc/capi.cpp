// capi.cpp -- the C ABI of include/cgli.h: contexts, resident fields and
// the drop-in entry points.  Host-side orchestration only; every number is
// computed by the kernels in flux.hip, viscous.hip, implicit.hip, bc.hip.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "cgli_internal.hpp"

using namespace cgli;

namespace {

enum { T_FLUX_X, T_FLUX_Y, T_FLUX_Z, T_VISCOUS, T_RHS, T_FWD, T_BWD, T_UPDATE, T_COEF, T_FILTER, T_SCATTER, T_GATHER,
       T_NCLASS };

struct EventPair { hipEvent_t a, b; };

}  // namespace

struct cgli_graph {
    cgli_ctx *ctx = nullptr;
    hipGraphExec_t exec = nullptr;
    long generation = 0;         // of the context at capture
    bool norm = false;           // ends with a deferred residual norm
    bool norm_wave = false;
    bool scatters_nr = false;    // holds a schedule-3 scatter: NL is then to be copied from NR where read
    bool nr_march = false;       // holds a schedule-4 update: NL and NR are then in march order only
    bool u4_fresh = false;       // holds one, which leaves U_'s march-ordered copy current
};

struct cgli_ctx {
    cgli_config cfg;
    int device = 0;
    int cus = 256;               // compute units of the device
    hipStream_t stream = nullptr;
    Geo g;
    Phys p;
    double *f[CGLI_NFIELDS] = {};
    bool valid[CGLI_NFIELDS] = {};
    double *scratch = nullptr;   // reductions
    long scratch_len = 0;
    double *er_host = nullptr;   // pinned
    int *ijk_dev = nullptr;
    std::vector<int> Nijk;       // (gend+NY+NZ+8) x 2
    std::vector<int> ijk;        // (cells+1) x 3
    int diss_scale = 0;          // Roe dissipation scaling: 0 none, 1 scalar Ep, 2 filter sensor
    bool nl_from_nr = false;     // NL's interior equals NR's and is not written yet (schedule 3)
    long nr_scatters = 0;        // schedule-3 scatters queued (a graph that holds one sets nl_from_nr when launched)
    long nr_scatters_at_capture = 0;
    bool nr_march = false;       // NL's and NR's interiors are the last schedule-4 increments, held in march
                                 // order only (written to the fields where read)
    long nr_marches = 0;         // schedule-4 updates queued
    long nr_marches_at_capture = 0;
    bool u4_fresh = false;       // u4's cell slots hold U_'s interior as the last schedule-4 update wrote it
    bool capturing = false;      // between cgli_graph_begin and cgli_graph_end
    long generation = 0;         // bumped whenever a captured graph would go stale
    bool norm_pending = false;   // a deferred residual norm awaits cgli_dual_time_stepping_wait
    hipEvent_t norm_ev = nullptr;    // recorded after the norm's copy to the host
    double *er_mapped = nullptr;     // device address of er_host, where the runtime maps it
    int *flags_mapped = nullptr;     // device address of counters_host
    bool norm_wave = false;      // ... with the wavefront's timeout flag
    int lusgs_mode = 4;          // 0: one launch per hyperplane, 1: tiled wavefront, 3: tiled wavefront on
                                 // march-ordered records, 4: march computing its coefficients (default)
    int sweep_waves = 0;         // schedule 4: waves per tile (0: by tile count; 1: one wave; 2-4: a chain
                                 // wave and helpers)
    double *coef = nullptr;      // split-coefficient blocks of the wavefront (float in the mixed mode)
    double *dgn = nullptr;       // diagonal blocks of the wavefront (FP64)
    void *coefM = nullptr;       // the same records in march order (mode 3)
    double *dgnM = nullptr;
    double *marchF = nullptr;    // mode 3: MR, NL and NR in march order
    void *geo4 = nullptr;        // mode 4: J and the 9 metrics in march order
    void *u4 = nullptr;          // mode 4: U_ in march order
    void *x4 = nullptr;          // mode 4: MR, NL and NR in march order
    void *rec4 = nullptr;        // mode 4: tile-edge exchange records (j products, then k products)
    long rec4_bytes = 0;         // bytes of each of the two
    void *gcol4 = nullptr;       // mode 4: each tile lane's column geometry, when the metrics do not vary along i
    bool geo_cols = false;       // the metrics do not vary along i (as of geo4_version)
    long geo_version = 0;        // bumped when J or a cell metric is written
    long geo4_version = -1;      // geo_version of geo4's contents
    bool geo_volatile = false;   // a device pointer to J or a metric was handed out
    int *morder_dev = nullptr;   // mode 3: tile order of its tiles
    int mtiles = 0;
    double *xbuf = nullptr;      // tile-edge exchange of the wavefront: 5 values per cell
    double *trash = nullptr;     // write-only scratch of the wavefront
    int *order_dev = nullptr;    // wavefront tile order
    int ntiles = 0;
    int *counters = nullptr;     // ticket, error flag
    double *filt = nullptr;      // filter work arrays (allocated on first use)
    double *prof = nullptr;      // 11 x ny wall-normal profiles
    int last_fp32 = 0;           // FP32-residual inner iterations of the last channel step
    ChannelState chan;           // rule, guess and network of the channel driver
    int *counters_host = nullptr;
    int max_spin = 1000000;
    bool timing = false;
    unsigned tmask = ~0u;        // kernel classes timed while timing is on
    std::vector<EventPair> ev[T_NCLASS];
    size_t ev_used[T_NCLASS] = {};
};

#define HIPCHK(x)                                      \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) return CGLI_E_HIP;       \
    } while (0)

// Scratch of the allocation-free entry point cgli_dplusgs_noalloc: the
// residual norm's partial sums (5 per block of 256 cells) and, in its last
// 8 doubles, the five norms.  Each device holds its own copy for the life of
// the process; every call on a device shares it, so calls that overlap on
// one device race.  Host code never passes the array's name to a kernel (it
// is a host-side stand-in there): it uses the device address noalloc_ptr.
constexpr long NOALLOC_DOUBLES = 1L << 19;   // 4 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static double *noalloc_ptr[NOALLOC_MAX_DEVICES];
static bool noalloc_queried[NOALLOC_MAX_DEVICES];

// once per device, with that device current (the first cgli_create on it)
static void noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_queried[device]) return;
    noalloc_queried[device] = true;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) noalloc_ptr[device] = static_cast<double *>(ptr);
    else (void)hipGetLastError();
}

static double *noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    return noalloc_ptr[device];
}

static int ensure(cgli_ctx *c, int id)
{
    if (c->f[id]) return CGLI_OK;
    if (c->capturing) return CGLI_E_STATE;   // allocations happen on an eager call first
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * c->g.N) != hipSuccess) return CGLI_E_NOMEM;
    if (hipMemsetAsync(p, 0, sizeof(double) * c->g.N, c->stream) != hipSuccess) return CGLI_E_HIP;
    c->f[id] = static_cast<double *>(p);
    return CGLI_OK;
}

static int ensure_range(cgli_ctx *c, int id, int n)
{
    for (int q = 0; q < n; q++) {
        int r = ensure(c, id + q);
        if (r) return r;
    }
    return CGLI_OK;
}

static bool have(cgli_ctx *c, int id, int n)
{
    for (int q = 0; q < n; q++)
        if (!c->valid[id + q]) return false;
    return true;
}

static void mark(cgli_ctx *c, int id, int n)
{
    for (int q = 0; q < n; q++) c->valid[id + q] = true;
}

// NL = NR on the interior, written where NL is read (schedule 3 leaves it
// to this: its own sweeps never read the field)
static void nl_sync(cgli_ctx *c)
{
    if (c->nr_march) {
        // schedule 4: the increments go from march order to NL and NR
        c->nr_march = false;
        c->nl_from_nr = false;
        launch_scatter4(c->g, static_cast<char *>(c->x4) + 2 * 16 * march4_field_pairs(c->g),
                        F5{{c->f[CGLI_NL], c->f[CGLI_NL + 1], c->f[CGLI_NL + 2], c->f[CGLI_NL + 3], c->f[CGLI_NL + 4]}},
                        F5{{c->f[CGLI_NR], c->f[CGLI_NR + 1], c->f[CGLI_NR + 2], c->f[CGLI_NR + 3], c->f[CGLI_NR + 4]}},
                        c->stream);
        return;
    }
    if (!c->nl_from_nr) return;
    c->nl_from_nr = false;
    launch_copy_interior(c->g, C5{{c->f[CGLI_NR], c->f[CGLI_NR + 1], c->f[CGLI_NR + 2], c->f[CGLI_NR + 3], c->f[CGLI_NR + 4]}},
                         F5{{c->f[CGLI_NL], c->f[CGLI_NL + 1], c->f[CGLI_NL + 2], c->f[CGLI_NL + 3], c->f[CGLI_NL + 4]}},
                         c->stream);
}

// J or a cell metric written: the march-ordered copy of the geometry is stale
static void geo_written(cgli_ctx *c, int id, int n)
{
    if (id <= CGLI_MET + 8 && id + n > CGLI_J) {
        c->geo_version++;
        c->generation++;   // a captured gather of the old geometry is stale
    }
}

static F5 f5(cgli_ctx *c, int id) { return F5{{c->f[id], c->f[id + 1], c->f[id + 2], c->f[id + 3], c->f[id + 4]}}; }
static C5 c5(cgli_ctx *c, int id) { return C5{{c->f[id], c->f[id + 1], c->f[id + 2], c->f[id + 3], c->f[id + 4]}}; }
static M9 m9(cgli_ctx *c, int id)
{
    M9 m;
    for (int q = 0; q < 9; q++) m.m[q] = c->f[id + q];
    return m;
}

static int ensure_scratch(cgli_ctx *c, long n)
{
    if (c->scratch_len >= n) return CGLI_OK;
    if (c->capturing) return CGLI_E_STATE;
    c->generation++;   // graphs captured with the old buffer are stale
    if (c->scratch) hipFree(c->scratch);
    c->scratch = nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * n) != hipSuccess) return CGLI_E_NOMEM;
    c->scratch = static_cast<double *>(p);
    c->scratch_len = n;
    return CGLI_OK;
}

// Kernel timing covers eager launches only (no events inside a graph).
static bool timed(const cgli_ctx *c, int cls) { return c->timing && !c->capturing && ((c->tmask >> cls) & 1u); }

static void tbegin(cgli_ctx *c, int cls)
{
    if (!timed(c, cls)) return;
    auto &v = c->ev[cls];
    if (c->ev_used[cls] == v.size()) {
        EventPair e;
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
        v.push_back(e);
    }
    hipEventRecord(v[c->ev_used[cls]].a, c->stream);
}

static void tend(cgli_ctx *c, int cls)
{
    if (!timed(c, cls)) return;
    hipEventRecord(c->ev[cls][c->ev_used[cls]].b, c->stream);
    c->ev_used[cls]++;
}

namespace cgli {
int ctx_field(cgli_ctx *c, int id, double **dptr)
{
    int r = ensure(c, id);
    if (!r) *dptr = c->f[id];
    return r;
}
hipStream_t ctx_stream(cgli_ctx *c) { return c->stream; }
const cgli_config *ctx_config(cgli_ctx *c) { return &c->cfg; }

void ctx_set_fp32(cgli_ctx *c, bool on) { c->p.fp32 = on ? 1 : 0; }
bool ctx_timing_on(cgli_ctx *c) { return c->timing; }

void ctx_set_last_fp32(cgli_ctx *c, int n) { c->last_fp32 = n; }
void ctx_mark(cgli_ctx *c, int id, int n)
{
    mark(c, id, n);
    geo_written(c, id, n);
}
const Geo &ctx_geo(cgli_ctx *c) { return c->g; }
ChannelState &ctx_channel(cgli_ctx *c) { return c->chan; }
}  // namespace cgli

extern "C" {

void cgli_config_reference(cgli_config *c, int nx, int ny, int nz)
{
    std::memset(c, 0, sizeof(*c));
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->first = 1; c->last = 1;
    c->gstart = 0; c->gend0 = nx - 1; c->nx_global = nx;
    c->nx_inlet = 100; c->nx_outlet = 150;
    c->sponge = 1; c->xi_mode = 0; c->x_periodic = 0; c->precision = 0;
    c->dxi = 0.001; c->deta = 0.005; c->dzeta = 0.002;
    c->high = 0.05;
    c->K = 1.4; c->R = 287;
    c->Cv = 717.5; c->mu_L = 0.0000185; c->Pr_L = 0.71;
    c->suth_T = 298.0592; c->suth_S = 110.;
    c->rho_ref = 1.1842; c->grav = 9.81; c->f_const = 0.0; c->f_dplusgs = 0.5993;
    c->E_ref = 253250.0;
}

const char *cgli_error_string(int code)
{
    switch (code) {
    case CGLI_OK: return "ok";
    case CGLI_E_ARG: return "invalid argument";
    case CGLI_E_HIP: return "HIP runtime error";
    case CGLI_E_NOMEM: return "device allocation failed";
    case CGLI_E_NONFINITE: return "non-finite residual norm";
    case CGLI_E_STATE: return "a required input field was never set";
    case CGLI_E_NODEV: return "no HIP device";
    case CGLI_E_TIMEOUT: return "bounded device wait expired";
    case CGLI_E_IO: return "restart file could not be opened, read or written";
    case CGLI_E_SIZE: return "the grid exceeds the fixed scratch of the allocation-free entry point";
    default: return "unknown error";
    }
}

int cgli_create(const cgli_config *cfg, int device, cgli_ctx **out)
{
    if (!cfg || !out) return CGLI_E_ARG;
    *out = nullptr;
    // the reconstruction stencils need 5 cells per direction inside the slab
    if (cfg->nx < 4 || cfg->ny < 6 || cfg->nz < 4) return CGLI_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CGLI_E_NODEV;
    if (device < 0 || device >= ndev) return CGLI_E_ARG;
    cgli_ctx *c = new (std::nothrow) cgli_ctx();
    if (!c) return CGLI_E_NOMEM;
    c->cfg = *cfg;
    c->device = device;
    {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n > 0) c->cus = n;
        else (void)hipGetLastError();
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CGLI_E_HIP;
    }
    noalloc_setup(device);
    Geo &g = c->g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz;
    g.Ym = cfg->ny + 4; g.Zm = cfg->nz + 4;
    g.N = (long)(cfg->nx + 6) * g.Ym * g.Zm;
    g.gend = cfg->nx + 2; g.NY = cfg->ny + 1; g.NZ = cfg->nz + 1;
    g.NYY = g.NY + 1; g.NZZ = g.NZ + 1;
    g.first = cfg->first; g.last = cfg->last; g.xper = cfg->x_periodic;
    Phys &p = c->p;
    p.K = cfg->K; p.R = cfg->R; p.Cv = cfg->Cv; p.mu_L = cfg->mu_L; p.Pr_L = cfg->Pr_L;
    p.suth_T = cfg->suth_T; p.suth_S = cfg->suth_S;
    p.dxi = cfg->dxi; p.deta = cfg->deta; p.dzeta = cfg->dzeta;
    p.rho_ref = cfg->rho_ref; p.grav = cfg->grav; p.f_const = cfg->f_const; p.E_ref = cfg->E_ref;
    p.high = cfg->high;
    p.fp32 = cfg->precision == 1;
    void *h = nullptr;
    if (hipHostMalloc(&h, sizeof(double) * 8) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return CGLI_E_HIP;
    }
    c->er_host = static_cast<double *>(h);
    {
        void *d = nullptr;   // without a mapping the norm is copied instead
        if (hipHostGetDevicePointer(&d, h, 0) == hipSuccess) c->er_mapped = static_cast<double *>(d);
        else (void)hipGetLastError();
    }
    if (hipEventCreateWithFlags(&c->norm_ev, hipEventDisableTiming) != hipSuccess) {
        hipHostFree(h);
        hipStreamDestroy(c->stream);
        delete c;
        return CGLI_E_HIP;
    }
    *out = c;
    return CGLI_OK;
}

int cgli_destroy(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    channel_drop_graphs(c->chan);
    for (int q = 0; q < CGLI_NFIELDS; q++)
        if (c->f[q]) hipFree(c->f[q]);
    if (c->scratch) hipFree(c->scratch);
    if (c->ijk_dev) hipFree(c->ijk_dev);
    if (c->coef) hipFree(c->coef);
    if (c->dgn) hipFree(c->dgn);
    if (c->coefM) hipFree(c->coefM);
    if (c->dgnM) hipFree(c->dgnM);
    if (c->marchF) hipFree(c->marchF);
    if (c->geo4) hipFree(c->geo4);
    if (c->u4) hipFree(c->u4);
    if (c->x4) hipFree(c->x4);
    if (c->rec4) hipFree(c->rec4);
    if (c->gcol4) hipFree(c->gcol4);
    if (c->morder_dev) hipFree(c->morder_dev);
    if (c->xbuf) hipFree(c->xbuf);
    if (c->trash) hipFree(c->trash);
    if (c->filt) hipFree(c->filt);
    if (c->prof) hipFree(c->prof);
    if (c->order_dev) hipFree(c->order_dev);
    if (c->counters) hipFree(c->counters);
    if (c->counters_host) hipHostFree(c->counters_host);
    if (c->er_host) hipHostFree(c->er_host);
    if (c->norm_ev) hipEventDestroy(c->norm_ev);
    for (int k = 0; k < T_NCLASS; k++)
        for (auto &e : c->ev[k]) {
            hipEventDestroy(e.a);
            hipEventDestroy(e.b);
        }
    hipStreamDestroy(c->stream);
    delete c;
    return CGLI_OK;
}

// Entry points that wait for the device cannot be captured into a graph.
#define NO_CAPTURE(c)                            \
    do {                                         \
        if ((c)->capturing) return CGLI_E_STATE; \
    } while (0)

int cgli_set_field(cgli_ctx *c, int id, const double *host)
{
    if (!c || !host || id < 0 || id >= CGLI_NFIELDS) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipSetDevice(c->device));
    int r = ensure(c, id);
    if (r) return r;
    // NL still to be copied from NR: copy it before either is overwritten
    if (id >= CGLI_NL && id < CGLI_NR + 5) nl_sync(c);
    HIPCHK(hipMemcpyAsync(c->f[id], host, sizeof(double) * c->g.N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->valid[id] = true;
    geo_written(c, id, 1);
    return CGLI_OK;
}

int cgli_get_field(cgli_ctx *c, int id, double *host)
{
    if (!c || !host || id < 0 || id >= CGLI_NFIELDS) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipSetDevice(c->device));
    int r = ensure(c, id);
    if (r) return r;
    if (id >= CGLI_NL && id < CGLI_NR + 5) nl_sync(c);   // NL, or both, may be held elsewhere
    HIPCHK(hipMemcpyAsync(host, c->f[id], sizeof(double) * c->g.N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CGLI_OK;
}

int cgli_field_device_ptr(cgli_ctx *c, int id, void **dptr)
{
    if (!c || !dptr || id < 0 || id >= CGLI_NFIELDS) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    // the deferred copy of NR into NL cannot be settled inside a capture
    if (c->capturing && id >= CGLI_NL && id < CGLI_NR + 5) return CGLI_E_STATE;
    int r = ensure(c, id);
    if (r) return r;
    if (id >= CGLI_NL && id < CGLI_NR + 5) nl_sync(c);
    c->valid[id] = true;   // the caller writes it
    if (id == CGLI_J || (id >= CGLI_MET && id < CGLI_MET + 9)) c->geo_volatile = true;
    *dptr = c->f[id];
    return CGLI_OK;
}

int cgli_device_memory(int device, size_t *free_bytes, size_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return CGLI_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CGLI_E_NODEV;
    if (device < 0 || device >= ndev) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
    return CGLI_OK;
}

int cgli_stream_copy(int device, size_t bytes, int reps, double ms[2])
{
    if (!ms || reps < 1 || bytes < 4096) return CGLI_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CGLI_E_NODEV;
    if (device < 0 || device >= ndev) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(device));
    const long n = (long)(bytes / 16) * 2;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double)) != hipSuccess) return CGLI_E_NOMEM;
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
        (void)hipFree(a);
        return CGLI_E_NOMEM;
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CGLI_OK;
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = CGLI_E_HIP;
    if (!rc) {
        launch_zero(a, n, s);
        // [0]: one double per lane (the field copy of the time shift);
        // [1]: 16 bytes per lane, grid-stride over 2048 blocks
        for (int v = 0; v < 2 && !rc; v++) {
            auto run = [&] {
                if (v == 0) launch_copy_field(b, a, n, s);
                else launch_stream_copy(b, a, n, 2048, s);
            };
            run();
            hipEventRecord(e0, s);
            for (int r = 0; r < reps; r++) run();
            hipEventRecord(e1, s);
            float t = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                rc = CGLI_E_HIP;
            ms[v] = (double)t / reps;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int cgli_synchronize(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    return CGLI_OK;
}

static int flux_common(cgli_ctx *c, int dir, double Ep, double e)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    const int jf = dir == 0 ? CGLI_J_U : dir == 1 ? CGLI_J_V : CGLI_J_W;
    const int mf = jf + 1;
    if (!have(c, CGLI_U_, 5) || !have(c, CGLI_J, 1) || !have(c, jf, 1) || !have(c, mf + 3 * dir, 3))
        return CGLI_E_STATE;
    const int out = dir == 0 ? CGLI_INFX : dir == 1 ? CGLI_INFY : CGLI_INFZ;
    int r = ensure_range(c, out, 5);
    if (r) return r;
    r = ensure_range(c, mf, 9);   // unset metric components are never read
    if (r) return r;
    const int ep_id = CGLI_EPX + dir;
    if (c->diss_scale == 2 && !have(c, ep_id, 1)) return CGLI_E_STATE;
    const double epS = c->diss_scale == 1 ? Ep : 1.0;
    const double *EpF = c->diss_scale == 2 ? c->f[ep_id] : nullptr;
    tbegin(c, T_FLUX_X + dir);
    if (dir == 0)
        launch_flux_x(c->g, c->p, e, c5(c, CGLI_U_), c->f[CGLI_J], c->f[jf], m9(c, mf), f5(c, out), epS, EpF,
                      c->stream);
    else if (dir == 1)
        launch_flux_y(c->g, c->p, Ep, e, c->cfg.xi_mode, c5(c, CGLI_U_), c->f[CGLI_J], c->f[jf], m9(c, mf),
                      f5(c, out), epS, EpF, c->stream);
    else
        launch_flux_z(c->g, c->p, e, c->cfg.xi_mode, c5(c, CGLI_U_), c->f[CGLI_J], c->f[jf], m9(c, mf),
                      f5(c, out), epS, EpF, c->stream);
    tend(c, T_FLUX_X + dir);
    HIPCHK(hipGetLastError());
    mark(c, out, 5);
    return CGLI_OK;
}

int cgli_flux_x(cgli_ctx *c, double Ep, double e) { return flux_common(c, 0, Ep, e); }
int cgli_flux_y(cgli_ctx *c, double Ep, double e) { return flux_common(c, 1, Ep, e); }
int cgli_flux_z(cgli_ctx *c, double Ep, double e) { return flux_common(c, 2, Ep, e); }

int cgli_set_dissipation_scaling(cgli_ctx *c, int mode)
{
    if (!c || mode < 0 || mode > 2) return CGLI_E_ARG;
    NO_CAPTURE(c);
    c->diss_scale = mode;
    c->generation++;
    return CGLI_OK;
}

int cgli_viscous(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (!have(c, CGLI_U_, 5) || !have(c, CGLI_J, 1) || !have(c, CGLI_J_U, 10) || !have(c, CGLI_J_V, 10) ||
        !have(c, CGLI_J_W, 10))
        return CGLI_E_STATE;
    int r = ensure_range(c, CGLI_VF, 4);
    if (r) return r;
    F4 vF{{c->f[CGLI_VF], c->f[CGLI_VF + 1], c->f[CGLI_VF + 2], c->f[CGLI_VF + 3]}};
    tbegin(c, T_VISCOUS);
    launch_viscous(c->g, c->p, c5(c, CGLI_U_), c->f[CGLI_J], c->f[CGLI_J_U], m9(c, CGLI_MET_U), c->f[CGLI_J_V],
                   m9(c, CGLI_MET_V), c->f[CGLI_J_W], m9(c, CGLI_MET_W), vF, c->stream);
    tend(c, T_VISCOUS);
    HIPCHK(hipGetLastError());
    mark(c, CGLI_VF, 4);
    return CGLI_OK;
}

int cgli_filter(cgli_ctx *c, double Roe_criterion, double E_high, double E_low)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (!have(c, CGLI_U_, 4) || !have(c, CGLI_J, 1) || !have(c, CGLI_YPOINT, 1)) return CGLI_E_STATE;
    if (c->g.nx < 3) return CGLI_E_ARG;
    int r = ensure_range(c, CGLI_EPX, 3);
    if (r) return r;
    if (!c->filt && c->capturing) return CGLI_E_STATE;
    if (!c->filt) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * filter_scratch_doubles(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->filt = static_cast<double *>(p);
        // NaN (FP64 and FP32 alike): a work cell read before the passes
        // write it would spoil the sensor at once
        HIPCHK(hipMemsetAsync(c->filt, 0xFF, sizeof(double) * filter_scratch_doubles(c->g), c->stream));
    }
    tbegin(c, T_FILTER);
    launch_filter(c->g, Roe_criterion, E_high, E_low, c5(c, CGLI_U_), c->f[CGLI_J], c->f[CGLI_YPOINT], c->filt,
                  c->cfg.precision == 1, c->f[CGLI_EPX], c->f[CGLI_EPY], c->f[CGLI_EPZ], c->stream);
    tend(c, T_FILTER);
    HIPCHK(hipGetLastError());
    mark(c, CGLI_EPX, 3);
    return CGLI_OK;
}

int cgli_filter_energies(cgli_ctx *c, double *Eh, double *E2h)
{
    if (!c || !Eh || !E2h) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipSetDevice(c->device));
    if (!c->filt || !c->valid[CGLI_EPX]) return CGLI_E_STATE;
    const long N = c->g.N;
    // the passes' work set ends with E_h and E_2h (FP32 in the mixed mode)
    if (c->cfg.precision == 1) {
        const float *w = reinterpret_cast<const float *>(c->filt);
        std::vector<float> tmp(2 * N);
        HIPCHK(hipMemcpyAsync(tmp.data(), w + 24 * N, sizeof(float) * 2 * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (long q = 0; q < N; q++) {
            Eh[q] = tmp[q];
            E2h[q] = tmp[N + q];
        }
    } else {
        HIPCHK(hipMemcpyAsync(Eh, c->filt + 24 * N, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(E2h, c->filt + 25 * N, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return CGLI_OK;
}

int cgli_channel_last_step(cgli_ctx *c, int *inner_fp32)
{
    if (!c || !inner_fp32) return CGLI_E_ARG;
    *inner_fp32 = c->last_fp32;
    return CGLI_OK;
}

int cgli_profiles(cgli_ctx *c, double *out)
{
    if (!c || !out) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipSetDevice(c->device));
    if (!have(c, CGLI_U_, 5) || !have(c, CGLI_J, 1)) return CGLI_E_STATE;
    const size_t bytes = sizeof(double) * 11 * c->g.ny;
    if (!c->prof) {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return CGLI_E_NOMEM;
        c->prof = static_cast<double *>(p);
    }
    launch_profiles(c->g, c->p, c5(c, CGLI_U_), c->f[CGLI_J], c->prof, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->prof, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CGLI_OK;
}

// Hyperplane.cpp:38-78 ordering: planes m ascending; inside a plane i, j, k
// ascending.  Built in O(cells) and uploaded once.
static int build_planes(cgli_ctx *c)
{
    if (c->ijk_dev) return CGLI_OK;
    const Geo &g = c->g;
    const int rows = g.gend + g.NY + g.NZ + 8;
    const long cells = (long)g.nx * g.ny * g.nz;
    c->Nijk.assign(2 * rows, 0);
    c->ijk.assign(3 * (cells + 1), 0);
    int ii = 0;
    for (int m = 7; m <= g.gend + g.NY + g.NZ; m++) {
        int mi = 0;
        for (int i = 3; i <= g.gend; i++)
            for (int j = 2; j <= g.NY; j++) {
                const int k = m - i - j;
                if (k < 2 || k > g.NZ) continue;
                mi++;
                ii++;
                c->ijk[3 * ii] = i;
                c->ijk[3 * ii + 1] = j;
                c->ijk[3 * ii + 2] = k;
            }
        c->Nijk[2 * m] = ii;
        c->Nijk[2 * m + 1] = mi;
    }
    if (c->capturing) return CGLI_E_STATE;
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(int) * c->ijk.size()) != hipSuccess) return CGLI_E_NOMEM;
    c->ijk_dev = static_cast<int *>(p);
    HIPCHK(hipMemcpy(c->ijk_dev, c->ijk.data(), sizeof(int) * c->ijk.size(), hipMemcpyHostToDevice));
    return CGLI_OK;
}

int cgli_hyperplane(cgli_ctx *c, int *Nijk, int *ijk)
{
    if (!c || !Nijk || !ijk) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    int r = build_planes(c);
    if (r) return r;
    std::memcpy(Nijk, c->Nijk.data(), sizeof(int) * c->Nijk.size());
    std::memcpy(ijk, c->ijk.data(), sizeof(int) * c->ijk.size());
    return CGLI_OK;
}

static int implicit_inputs(cgli_ctx *c)
{
    if (!have(c, CGLI_U_, 5) || !have(c, CGLI_U, 5) || !have(c, CGLI_UQ, 5) || !have(c, CGLI_J, 1) ||
        !have(c, CGLI_MET, 9) || !have(c, CGLI_INFX, 15) || !have(c, CGLI_VF, 4))
        return CGLI_E_STATE;
    int r = ensure_range(c, CGLI_MR, 15);
    if (r) return r;
    return ensure_range(c, CGLI_RES, 5);
}

// the norm's last reduction and its copies to pinned host memory, queued
static int enqueue_norm(cgli_ctx *c, int nb, bool check_wave)
{
    const double DN = 1. / (double)((long)c->cfg.nx_global * c->cfg.ny * c->cfg.nz);
    if (c->er_mapped && (!check_wave || c->flags_mapped)) {
        // the last reduction writes the norms (and the wavefront's flags)
        // into pinned host memory itself: no copy waits in the stream
        launch_finish_norm(nb, c->scratch, DN, 0, c->er_mapped, c->stream, check_wave ? c->counters : nullptr,
                           c->flags_mapped);
        HIPCHK(hipGetLastError());
    } else {
        double *er_dev = c->scratch + c->scratch_len - 8;
        launch_finish_norm(nb, c->scratch, DN, 0, er_dev, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->er_host, er_dev, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
        if (check_wave)
            HIPCHK(hipMemcpyAsync(c->counters_host, c->counters, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
    }
    // (a graph launch records the event after the graph instead)
    if (!c->capturing) HIPCHK(hipEventRecord(c->norm_ev, c->stream));
    return CGLI_OK;
}

// wait for the queued norm (not for work queued after it) and check it
static int collect_norm(cgli_ctx *c, double er[5], bool check_wave)
{
    HIPCHK(hipEventSynchronize(c->norm_ev));
    if (check_wave && c->counters_host[1] != 0) return CGLI_E_TIMEOUT;
    bool finite = true;
    for (int v = 0; v < 5; v++) {
        er[v] = c->er_host[v];
        if (!std::isfinite(er[v])) finite = false;
    }
    return finite ? CGLI_OK : CGLI_E_NONFINITE;
}

static int finish_norm(cgli_ctx *c, int nb, double er[5], bool check_wave = false)
{
    int r = enqueue_norm(c, nb, check_wave);
    return r ? r : collect_norm(c, er, check_wave);
}

// schedule 4's waves per tile: by default a chain wave with two helpers
// when there are at most two tiles per CU (the chain wave's instruction
// stream bounds the sweep, and the helpers find idle SIMDs: 128^3 forward
// sweep 0.34 ms against 0.47 with one wave, 192x128x192 0.63 against 0.66),
// one wave otherwise (the helpers then share SIMDs with chain waves:
// 256x192x256 1.16 ms against 1.06)
static int sweep_waves(const cgli_ctx *c)
{
    return c->sweep_waves ? c->sweep_waves : (c->mtiles <= 2 * c->cus ? 3 : 1);
}

// the wavefront's records (in the layout of the current schedule) and its
// exchange buffers, allocated on first use
static int wave_setup(cgli_ctx *c)
{
    void *p = nullptr;
    if (c->lusgs_mode == 3 && !march_fits(c->g, c->cfg.precision == 1)) return CGLI_E_ARG;
    if (c->capturing && ((c->lusgs_mode == 3 && !c->coefM) || (c->lusgs_mode == 4 && !c->geo4) ||
                         (c->lusgs_mode < 3 && !c->coef) || !c->xbuf))
        return CGLI_E_STATE;
    if (c->lusgs_mode >= 3 && !c->morder_dev) {
        std::vector<int> order;
        c->mtiles = wave_tiles(c->g, MARCH_TJ, MARCH_TK, order);
        if (hipMalloc(&p, sizeof(int) * order.size()) != hipSuccess) return CGLI_E_NOMEM;
        c->morder_dev = static_cast<int *>(p);
        HIPCHK(hipMemcpy(c->morder_dev, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    }
    if (c->lusgs_mode == 4 && !c->geo4) {
        const long gb = 16 * march4_geo_pairs(c->g), fb = 16 * march4_field_pairs(c->g);
        if (hipMalloc(&p, gb) != hipSuccess) return CGLI_E_NOMEM;
        c->geo4 = p;
        if (hipMalloc(&p, fb) != hipSuccess) return CGLI_E_NOMEM;
        c->u4 = p;
        c->u4_fresh = false;
        if (hipMalloc(&p, 3 * fb) != hipSuccess) return CGLI_E_NOMEM;
        c->x4 = p;
        // slots with no cell are never written: NaN in them
        HIPCHK(hipMemset(c->geo4, 0xFF, gb));
        HIPCHK(hipMemset(c->u4, 0xFF, fb));
        HIPCHK(hipMemset(c->x4, 0xFF, 3 * fb));
        c->geo4_version = -1;
        // the exchange records: every offset a sweep forms stays below 2 GiB
        c->rec4_bytes = march4_rec_bytes(c->g);
        if (2 * c->rec4_bytes + 64 > (1L << 31)) return CGLI_E_ARG;
        if (hipMalloc(&p, 2 * c->rec4_bytes) != hipSuccess) return CGLI_E_NOMEM;
        c->rec4 = p;
        if (hipMalloc(&p, 16 * march4_col_pairs(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->gcol4 = p;
    }
    if (c->lusgs_mode == 3 && !c->coefM) {
        // FP64 pairs are 16 bytes: the FP32 records use half
        if (hipMalloc(&p, 16 * march_coef_pairs(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->coefM = p;
        if (hipMalloc(&p, 16 * march_diag_pairs(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->dgnM = static_cast<double *>(p);
        if (hipMalloc(&p, sizeof(double) * 3 * march_field_doubles(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->marchF = static_cast<double *>(p);
        // slots with no cell are never written: NaN in them, so that a read
        // of one where a cell's value belongs spoils every result at once
        HIPCHK(hipMemset(c->coefM, 0xFF, 16 * march_coef_pairs(c->g)));
        HIPCHK(hipMemset(c->dgnM, 0xFF, 16 * march_diag_pairs(c->g)));
        HIPCHK(hipMemset(c->marchF, 0xFF, sizeof(double) * 3 * march_field_doubles(c->g)));
    }
    if (c->lusgs_mode < 3 && !c->coef) {
        if (hipMalloc(&p, sizeof(double) * coef_doubles(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->coef = static_cast<double *>(p);
        if (hipMalloc(&p, sizeof(double) * diag_doubles(c->g)) != hipSuccess) return CGLI_E_NOMEM;
        c->dgn = static_cast<double *>(p);
    }
    if (c->xbuf) return CGLI_OK;
    if (hipMalloc(&p, sizeof(double) * 5 * c->g.N) != hipSuccess) return CGLI_E_NOMEM;
    c->xbuf = static_cast<double *>(p);
    HIPCHK(hipMemset(c->xbuf, 0, sizeof(double) * 5 * c->g.N));   // ghost entries stay zero
    std::vector<int> order;
    c->ntiles = wave_tiles(c->g, WAVE_TJ, WAVE_TK, order);
    {
        // one 256-double scratch line per workgroup of either wavefront
        std::vector<int> morder;
        const int mt = wave_tiles(c->g, MARCH_TJ, MARCH_TK, morder);
        if (hipMalloc(&p, sizeof(double) * 256 * std::max(c->ntiles, mt)) != hipSuccess) return CGLI_E_NOMEM;
    }
    c->trash = static_cast<double *>(p);
    if (hipMalloc(&p, sizeof(int) * order.size()) != hipSuccess) return CGLI_E_NOMEM;
    c->order_dev = static_cast<int *>(p);
    HIPCHK(hipMemcpy(c->order_dev, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    if (hipMalloc(&p, sizeof(int) * 4) != hipSuccess) return CGLI_E_NOMEM;
    c->counters = static_cast<int *>(p);
    if (hipHostMalloc(&p, sizeof(int) * 4) != hipSuccess) return CGLI_E_NOMEM;
    c->counters_host = static_cast<int *>(p);
    {
        void *d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) == hipSuccess) c->flags_mapped = static_cast<int *>(d);
        else (void)hipGetLastError();
    }
    return CGLI_OK;
}

int cgli_set_lusgs_mode(cgli_ctx *c, int mode)
{
    if (!c || mode < 0 || mode > 4 || mode == 2) return CGLI_E_ARG;
    NO_CAPTURE(c);
    c->lusgs_mode = mode;
    c->generation++;
    return CGLI_OK;
}

int cgli_set_sweep_waves(cgli_ctx *c, int waves)
{
    if (!c || waves < 0 || waves > 4) return CGLI_E_ARG;
    NO_CAPTURE(c);
    c->sweep_waves = waves;
    c->generation++;
    return CGLI_OK;
}

int cgli_set_spin_limit(cgli_ctx *c, int max_spin)
{
    if (!c || max_spin < 1) return CGLI_E_ARG;
    NO_CAPTURE(c);
    c->max_spin = max_spin;
    c->generation++;
    return CGLI_OK;
}

int cgli_dual_time_stepping(cgli_ctx *c, double deltaT, double e, double er[5], int sweeps, int write_residual)
{
    if (c) NO_CAPTURE(c);
    return ctx_dts(c, deltaT, e, er, sweeps, write_residual, false);
}

int cgli_dual_time_stepping_async(cgli_ctx *c, double deltaT, double e, int sweeps, int write_residual)
{
    return ctx_dts_async(c, deltaT, e, sweeps, write_residual, false);
}

int cgli_dual_time_stepping_wait(cgli_ctx *c, double er[5])
{
    if (!c || !er) return CGLI_E_ARG;
    NO_CAPTURE(c);
    return ctx_dts_wait(c, er);
}

int cgli_graph_begin(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    if (c->capturing) return CGLI_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    c->norm_pending = false;
    c->nr_scatters_at_capture = c->nr_scatters;
    c->nr_marches_at_capture = c->nr_marches;
    return CGLI_OK;
}

int cgli_graph_end(cgli_ctx *c, cgli_graph **out)
{
    if (!c || !out) return CGLI_E_ARG;
    *out = nullptr;
    if (!c->capturing) return CGLI_E_STATE;
    c->capturing = false;
    const bool norm = c->norm_pending, wave = c->norm_wave;
    c->norm_pending = false;   // captured, not queued
    hipGraph_t gr = nullptr;
    if (hipStreamEndCapture(c->stream, &gr) != hipSuccess || !gr) return CGLI_E_HIP;
    cgli_graph *G = new (std::nothrow) cgli_graph();
    if (!G) {
        hipGraphDestroy(gr);
        return CGLI_E_NOMEM;
    }
    const hipError_t e = hipGraphInstantiate(&G->exec, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (e != hipSuccess) {
        delete G;
        return CGLI_E_HIP;
    }
    G->ctx = c;
    G->generation = c->generation;
    G->norm = norm;
    G->norm_wave = wave;
    G->scatters_nr = c->nr_scatters != c->nr_scatters_at_capture;
    G->nr_march = c->nr_marches != c->nr_marches_at_capture;
    G->u4_fresh = G->nr_march;
    *out = G;
    return CGLI_OK;
}

int cgli_graph_launch(cgli_graph *G)
{
    if (!G || !G->ctx) return CGLI_E_ARG;
    cgli_ctx *c = G->ctx;
    if (c->capturing || G->generation != c->generation) return CGLI_E_STATE;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipGraphLaunch(G->exec, c->stream));
    if (G->scatters_nr) c->nl_from_nr = true;
    if (G->nr_march) {
        c->nr_march = true;
        c->nl_from_nr = false;
    }
    if (G->u4_fresh) c->u4_fresh = true;
    if (G->norm) {
        HIPCHK(hipEventRecord(c->norm_ev, c->stream));
        c->norm_pending = true;
        c->norm_wave = G->norm_wave;
    }
    return CGLI_OK;
}

int cgli_graph_destroy(cgli_graph *G)
{
    if (!G) return CGLI_E_ARG;
    if (G->exec) hipGraphExecDestroy(G->exec);
    delete G;
    return CGLI_OK;
}

}  // extern "C"

namespace cgli {

// rhs_ready: MR, NL, NR already hold rhs_diag's output for the current U_
// (ctx_rhs_sumsq was the last call on it), so rhs_diag is not repeated
int ctx_dts(cgli_ctx *c, double deltaT, double e, double er[5], int sweeps, int write_residual, bool rhs_ready)
{
    if (!c || !er) return CGLI_E_ARG;
    int r = ctx_dts_async(c, deltaT, e, sweeps, write_residual, rhs_ready);
    return r ? r : ctx_dts_wait(c, er);
}

int ctx_dts_wait(cgli_ctx *c, double er[5])
{
    if (!c->norm_pending) return CGLI_E_STATE;
    c->norm_pending = false;
    return collect_norm(c, er, c->norm_wave);
}

int ctx_dts_async(cgli_ctx *c, double deltaT, double e, int sweeps, int write_residual, bool rhs_ready,
                  bool u4_current)
{
    if (!c || sweeps < 1) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    int r = implicit_inputs(c);
    if (r) return r;
    r = build_planes(c);
    if (r) return r;
    const Geo &g = c->g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const long nb = std::max((ncell + 255) / 256, c->lusgs_mode == 4 ? update4_blocks(g) : 0L);
    r = ensure_scratch(c, 5 * nb + (long)g.NYY * g.NZZ + 64);
    if (r) return r;
    const hipStream_t s = c->stream;
    if (c->lusgs_mode >= 1) {
        r = wave_setup(c);
        if (r) return r;
    }
    // other schedules write NL and NR in place: bring the last schedule-4
    // increments to them first (this call may yet fail before writing)
    if (c->lusgs_mode != 4) nl_sync(c);
    // also zeroes the wavefront's ticket and error flag
    launch_zero_increment_ghosts(g, f5(c, CGLI_NL), f5(c, CGLI_NR), c->lusgs_mode >= 1 ? c->counters : nullptr, s);
    if (c->cfg.sponge || write_residual)
        for (int v = 0; v < 5; v++) launch_zero(c->f[CGLI_RES + v], g.N, s);
    if (c->cfg.sponge)
        launch_sponge(g, c->p, e, c->cfg.gstart, c->cfg.gend0, c->cfg.nx_global, c->cfg.nx_inlet,
                      c->cfg.nx_outlet, c5(c, CGLI_U_), c->f[CGLI_J], c->f[CGLI_MET + 4], f5(c, CGLI_RES),
                      c->scratch, s);
    if (!rhs_ready || c->cfg.sponge) {
        tbegin(c, T_RHS);
        launch_rhs_diag(g, c->p, deltaT, e, c->cfg.sponge, c5(c, CGLI_U_), c5(c, CGLI_U), c5(c, CGLI_UQ),
                        c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_INFX), c5(c, CGLI_INFY), c5(c, CGLI_INFZ),
                        C4{{c->f[CGLI_VF], c->f[CGLI_VF + 1], c->f[CGLI_VF + 2], c->f[CGLI_VF + 3]}},
                        c5(c, CGLI_RES), f5(c, CGLI_MR), f5(c, CGLI_NL), f5(c, CGLI_NR), s, c->lusgs_mode < 3);
        tend(c, T_RHS);
    }
    const int mlast = g.gend + g.NY + g.NZ;
    const long f4b = 16 * march4_field_pairs(g);
    char *const x4 = static_cast<char *>(c->x4);   // mode 4: MR, NL, NR in march order
    if (c->lusgs_mode == 4) {
        // the geometry in march order: per column when the metrics do not
        // vary along i (checked once per geometry version), per cell
        // otherwise, and every call while a caller holds a device pointer to
        // them (it may write them at any time)
        bool full_geo = false;
        if (c->geo_volatile) {
            full_geo = true;
            c->geo_cols = false;
        } else if (c->geo4_version != c->geo_version) {
            if (c->capturing) return CGLI_E_STATE;
            int *flag = c->counters + 2;
            HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), s));
            launch_geo_icheck(g, c->f[CGLI_J], m9(c, CGLI_MET), flag, s);
            int differs = 0;
            HIPCHK(hipMemcpyAsync(&differs, flag, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            c->geo_cols = differs == 0;
            if (c->geo_cols) launch_gather_cols4(g, c->f[CGLI_J], m9(c, CGLI_MET), c->gcol4, s);
            else full_geo = true;
            c->geo4_version = c->geo_version;
        }
        // U_'s march-ordered copy is still the one the last update wrote
        // when the caller vouches that U_'s interior has not changed since
        const bool keep_u4 = u4_current && c->u4_fresh;
        tbegin(c, T_GATHER);
        launch_gather4(g, c5(c, CGLI_U_), c5(c, CGLI_MR), c->f[CGLI_J], m9(c, CGLI_MET), keep_u4 ? nullptr : c->u4,
                       x4, full_geo ? c->geo4 : nullptr, s);
        tend(c, T_GATHER);
    } else if (c->lusgs_mode >= 1) {
        tbegin(c, T_COEF);
        if (c->lusgs_mode == 3)
            launch_split_march(g, c->p, deltaT, e, c5(c, CGLI_U_), c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_MR),
                               c->coefM, c->dgnM, c->marchF, c->cfg.precision == 1, s);
        else
            launch_split_coef(g, c->p, deltaT, e, c5(c, CGLI_U_), c->f[CGLI_J], m9(c, CGLI_MET), c->coef, c->dgn,
                              c->cfg.precision == 1, s);
        tend(c, T_COEF);
    }
    for (int sw = 0; sw < sweeps; sw++) {
        if (c->lusgs_mode == 4) {
            launch_rec4_prep(c->rec4, c->rec4_bytes, c->counters, s);
            tbegin(c, T_FWD);
            launch_lusgs_fused(g, c->p, deltaT, e, 1, c->geo4, c->geo_cols ? c->gcol4 : nullptr, c->u4, x4, x4 + f4b,
                               c5(c, CGLI_U_), c->f[CGLI_J],
                               m9(c, CGLI_MET), c5(c, CGLI_NL), c->rec4, c->rec4_bytes, c->morder_dev,
                               c->mtiles, c->counters, c->max_spin, -1, sweep_waves(c), s);
            tend(c, T_FWD);
            launch_rec4_prep(c->rec4, c->rec4_bytes, c->counters, s);
            tbegin(c, T_BWD);
            launch_lusgs_fused(g, c->p, deltaT, e, 0, c->geo4, c->geo_cols ? c->gcol4 : nullptr, c->u4, x4 + f4b,
                               x4 + 2 * f4b, c5(c, CGLI_U_),
                               c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_NR), c->rec4, c->rec4_bytes,
                               c->morder_dev, c->mtiles, c->counters, c->max_spin, -1, sweep_waves(c), s);
            tend(c, T_BWD);
            // the last sweep's increments stay in march order: the update
            // reads them there (k_update4)
        } else if (c->lusgs_mode == 3) {
            launch_march_prep(g, 1, c->xbuf, c->counters, s);
            tbegin(c, T_FWD);
            const long mf = march_field_doubles(g);
            launch_lusgs_march(g, 1, c->coefM, c->dgnM, c->cfg.precision == 1, c->marchF, c->marchF + mf,
                               c5(c, CGLI_NL), c->xbuf, c->trash, c->morder_dev, c->mtiles, c->counters, c->max_spin,
                               MARCH_SLACK, s);
            tend(c, T_FWD);
            launch_march_prep(g, 0, c->xbuf, c->counters, s);
            tbegin(c, T_BWD);
            launch_lusgs_march(g, 0, c->coefM, c->dgnM, c->cfg.precision == 1, c->marchF + mf,
                               c->marchF + 2 * mf, c5(c, CGLI_NR), c->xbuf, c->trash, c->morder_dev, c->mtiles,
                               c->counters, c->max_spin, MARCH_SLACK, s);
            tend(c, T_BWD);
            if (sw == sweeps - 1) {
                tbegin(c, T_SCATTER);
                launch_march_scatter(g, c->marchF + 2 * mf, F5{{nullptr, nullptr, nullptr, nullptr, nullptr}},
                                     f5(c, CGLI_NR), s);
                c->nl_from_nr = true;
                c->nr_scatters++;
                tend(c, T_SCATTER);
            }
        } else if (c->lusgs_mode >= 1) {
            launch_lusgs_wave_prep(g, c->xbuf, c->counters, s);
            tbegin(c, T_FWD);
            launch_lusgs_wave(g, 1, c->coef, c->dgn, c->cfg.precision == 1, c5(c, CGLI_MR), f5(c, CGLI_NL), c->xbuf, c->trash, c->order_dev,
                              c->ntiles, c->counters, c->max_spin, s);
            tend(c, T_FWD);
            launch_lusgs_wave_prep(g, c->xbuf, c->counters, s);
            tbegin(c, T_BWD);
            launch_lusgs_wave(g, 0, c->coef, c->dgn, c->cfg.precision == 1, c5(c, CGLI_NL), f5(c, CGLI_NR), c->xbuf, c->trash, c->order_dev,
                              c->ntiles, c->counters, c->max_spin, s);
            tend(c, T_BWD);
        } else {
            tbegin(c, T_FWD);
            for (int m = 7; m <= mlast; m++) {
                const int cnt = c->Nijk[2 * m + 1];
                if (!cnt) continue;
                launch_lusgs_plane(g, c->p, deltaT, e, 1, m, c->Nijk[2 * m] - cnt + 1, cnt, c->ijk_dev,
                                   c5(c, CGLI_U_), c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_MR), f5(c, CGLI_NL),
                                   f5(c, CGLI_NR), s);
            }
            tend(c, T_FWD);
            tbegin(c, T_BWD);
            for (int m = mlast; m >= 7; m--) {
                const int cnt = c->Nijk[2 * m + 1];
                if (!cnt) continue;
                launch_lusgs_plane(g, c->p, deltaT, e, 0, m, c->Nijk[2 * m] - cnt + 1, cnt, c->ijk_dev,
                                   c5(c, CGLI_U_), c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_MR), f5(c, CGLI_NL),
                                   f5(c, CGLI_NR), s);
            }
            tend(c, T_BWD);
        }
        // NL = NR at the end of a sweep pair (schedules 3, 4: where NL is read)
        if (c->lusgs_mode < 3) {
            launch_copy_interior(g, c5(c, CGLI_NR), f5(c, CGLI_NL), s);
            c->nl_from_nr = false;
        }
    }
    tbegin(c, T_UPDATE);
    int nblk;
    if (c->lusgs_mode == 4) {
        // U_ and NR in march order; NL and NR are written where read
        nblk = launch_update4(g, c->p, f5(c, CGLI_U_), c->f[CGLI_J], c->u4, x4 + 2 * f4b, f5(c, CGLI_RES),
                              write_residual, c->scratch, s);
        c->nr_march = true;
        c->nl_from_nr = false;
        c->nr_marches++;
        c->u4_fresh = true;
    } else {
        c->u4_fresh = false;
        nblk = launch_update_norm(g, c->p, f5(c, CGLI_U_), c->f[CGLI_J], c5(c, CGLI_NR), f5(c, CGLI_RES),
                                  write_residual, 0, c->scratch, s);
    }
    tend(c, T_UPDATE);
    HIPCHK(hipGetLastError());
    mark(c, CGLI_MR, 15);
    r = enqueue_norm(c, nblk, c->lusgs_mode >= 1);
    if (r) return r;
    c->norm_pending = true;
    c->norm_wave = c->lusgs_mode >= 1;
    return CGLI_OK;
}

}  // namespace cgli

extern "C" {

int cgli_dplusgs(cgli_ctx *c, double deltaT, double er[5], double e, int sweeps)
{
    if (!c || !er || sweeps < 1) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipSetDevice(c->device));
    int r = implicit_inputs(c);
    if (r) return r;
    const Geo &g = c->g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const long nb = (ncell + 255) / 256;
    r = ensure_scratch(c, 5 * nb + 64);
    if (r) return r;
    const hipStream_t s = c->stream;
    Phys pd = c->p;
    pd.grav = 0.0;            // DPLUSGS.cpp:417,420 use the constant f of prm.h:11
    pd.f_const = c->cfg.f_dplusgs;
    launch_zero_increment_ghosts(g, f5(c, CGLI_NL), f5(c, CGLI_NR), nullptr, s);
    launch_rhs_diag(g, pd, deltaT, e, 0, c5(c, CGLI_U_), c5(c, CGLI_U), c5(c, CGLI_UQ), c->f[CGLI_J],
                    m9(c, CGLI_MET), c5(c, CGLI_INFX), c5(c, CGLI_INFY), c5(c, CGLI_INFZ),
                    C4{{c->f[CGLI_VF], c->f[CGLI_VF + 1], c->f[CGLI_VF + 2], c->f[CGLI_VF + 3]}},
                    c5(c, CGLI_RES), f5(c, CGLI_MR), f5(c, CGLI_NL), f5(c, CGLI_NR), s);
    for (int sw = 0; sw < sweeps; sw++) {
        launch_dplur_sweep(g, c->p, deltaT, e, c5(c, CGLI_U_), c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_MR),
                           c5(c, CGLI_NL), f5(c, CGLI_NR), s);
        launch_copy_interior(g, c5(c, CGLI_NR), f5(c, CGLI_NL), s);
        c->nl_from_nr = false;
        c->nr_march = false;   // NL and NR now hold these increments
        c->u4_fresh = false;   // U_ changes outside the march order
    }
    for (int v = 0; v < 5; v++) launch_zero(c->f[CGLI_RES + v], g.N, s);
    const int nblk = launch_update_norm(g, c->p, f5(c, CGLI_U_), c->f[CGLI_J], c5(c, CGLI_NR), f5(c, CGLI_RES),
                                        1, 1, c->scratch, s);
    HIPCHK(hipGetLastError());
    mark(c, CGLI_MR, 15);
    return finish_norm(c, nblk, er);
}

int cgli_dplusgs_noalloc(cgli_ctx *c, double deltaT, double er[5], double e, int sweeps)
{
    if (!c || !er || sweeps < 1) return CGLI_E_ARG;
    NO_CAPTURE(c);
    const Geo &g = c->g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const long nb = (ncell + 255) / 256;
    if (5 * nb + 8 > NOALLOC_DOUBLES) return CGLI_E_SIZE;
    HIPCHK(hipSetDevice(c->device));
    double *ptr = noalloc_scratch(c->device);
    if (!ptr) return CGLI_E_HIP;
    // inputs and outputs already on the device: nothing is allocated here
    if (!have(c, CGLI_U_, 5) || !have(c, CGLI_U, 5) || !have(c, CGLI_UQ, 5) || !have(c, CGLI_J, 1) ||
        !have(c, CGLI_MET, 9) || !have(c, CGLI_INFX, 15) || !have(c, CGLI_VF, 4))
        return CGLI_E_STATE;
    for (int id = CGLI_MR; id < CGLI_RES + 5; id++)
        if (!c->f[id]) return CGLI_E_STATE;
    const hipStream_t s = c->stream;
    Phys pd = c->p;
    pd.grav = 0.0;            // DPLUSGS.cpp:417,420 use the constant f of prm.h:11
    pd.f_const = c->cfg.f_dplusgs;
    launch_zero_increment_ghosts(g, f5(c, CGLI_NL), f5(c, CGLI_NR), nullptr, s);
    launch_rhs_diag(g, pd, deltaT, e, 0, c5(c, CGLI_U_), c5(c, CGLI_U), c5(c, CGLI_UQ), c->f[CGLI_J],
                    m9(c, CGLI_MET), c5(c, CGLI_INFX), c5(c, CGLI_INFY), c5(c, CGLI_INFZ),
                    C4{{c->f[CGLI_VF], c->f[CGLI_VF + 1], c->f[CGLI_VF + 2], c->f[CGLI_VF + 3]}},
                    c5(c, CGLI_RES), f5(c, CGLI_MR), f5(c, CGLI_NL), f5(c, CGLI_NR), s);
    for (int sw = 0; sw < sweeps; sw++) {
        launch_dplur_sweep(g, c->p, deltaT, e, c5(c, CGLI_U_), c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_MR),
                           c5(c, CGLI_NL), f5(c, CGLI_NR), s);
        launch_copy_interior(g, c5(c, CGLI_NR), f5(c, CGLI_NL), s);
        c->nl_from_nr = false;
        c->nr_march = false;   // NL and NR now hold these increments
        c->u4_fresh = false;   // U_ changes outside the march order
    }
    for (int v = 0; v < 5; v++) launch_zero(c->f[CGLI_RES + v], g.N, s);
    // every partial the final reduction reads is written by this call
    const int nblk = launch_update_norm(g, c->p, f5(c, CGLI_U_), c->f[CGLI_J], c5(c, CGLI_NR), f5(c, CGLI_RES),
                                        1, 1, ptr, s);
    double *er_dev = ptr + NOALLOC_DOUBLES - 8;
    const double DN = 1. / (double)((long)c->cfg.nx_global * c->cfg.ny * c->cfg.nz);
    launch_finish_norm(nblk, ptr, DN, 0, er_dev, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->er_host, er_dev, sizeof(double) * 5, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    mark(c, CGLI_MR, 15);
    bool finite = true;
    for (int v = 0; v < 5; v++) {
        er[v] = c->er_host[v];
        if (!std::isfinite(er[v])) finite = false;
    }
    return finite ? CGLI_OK : CGLI_E_NONFINITE;
}

int cgli_bc_y(cgli_ctx *c, double T_low, double T_high, int heated_lo, int heated_hi)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (!have(c, CGLI_U_, 5) || !have(c, CGLI_J, 1)) return CGLI_E_STATE;
    launch_bc_y(c->g, c->p, c->cfg.gstart, f5(c, CGLI_U_), c->f[CGLI_J], T_low, T_high, heated_lo, heated_hi,
                c->stream);
    HIPCHK(hipGetLastError());
    return CGLI_OK;
}

int cgli_bc_z(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (!have(c, CGLI_U_, 5)) return CGLI_E_STATE;
    launch_bc_z(c->g, f5(c, CGLI_U_), c->stream);
    HIPCHK(hipGetLastError());
    return CGLI_OK;
}

int cgli_bc_x_periodic(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (!have(c, CGLI_U_, 5)) return CGLI_E_STATE;
    launch_bc_x_periodic(c->g, f5(c, CGLI_U_), c->stream);
    HIPCHK(hipGetLastError());
    return CGLI_OK;
}

}  // extern "C"

// Whole-buffer positional transfers (short counts from pread/pwrite repeat).
static bool file_xfer(int fd, void *buf, size_t bytes, off_t off, bool write)
{
    char *b = static_cast<char *>(buf);
    while (bytes) {
        const ssize_t n = write ? pwrite(fd, b, bytes, off) : pread(fd, b, bytes, off);
        if (n <= 0) return false;   // error, or end of file before the block ends
        b += n;
        off += n;
        bytes -= (size_t)n;
    }
    return true;
}

// The restart file of Output.cpp:167-284 / Initial_condition.cpp:112-219:
// block b (U1..U5, then U1q..U5q) holds nx_global*ny*nz doubles, and this
// slab's cells start gstart planes into it (the MPI_File_set_view offsets,
// Output.cpp:232-279).  One block at a time goes through a device staging
// buffer and a pinned host buffer.
static int restart_io(cgli_ctx *c, const char *path, bool write)
{
    if (!c || !path) return CGLI_E_ARG;
    NO_CAPTURE(c);
    HIPCHK(hipSetDevice(c->device));
    const cgli_config &cf = c->cfg;
    if (cf.gstart < 0 || cf.nx_global < cf.gstart + cf.nx) return CGLI_E_ARG;
    if (!have(c, CGLI_J, 1)) return CGLI_E_STATE;
    if (write && (!have(c, CGLI_U, 5) || !have(c, CGLI_UQ, 5))) return CGLI_E_STATE;
    if (!write) {
        int r = ensure_range(c, CGLI_U_, 15);   // U_, U, Uq
        if (r) return r;
    }
    const long plane = (long)cf.ny * cf.nz;
    const size_t bytes = sizeof(double) * (size_t)cf.nx * plane;
    void *dbuf = nullptr, *hbuf = nullptr;
    if (hipMalloc(&dbuf, bytes) != hipSuccess) return CGLI_E_NOMEM;
    if (hipHostMalloc(&hbuf, bytes) != hipSuccess) {
        hipFree(dbuf);
        return CGLI_E_NOMEM;
    }
    const int fd = write ? open(path, O_RDWR | O_CREAT, 0644) : open(path, O_RDONLY);
    int rc = fd < 0 ? CGLI_E_IO : CGLI_OK;
    for (int b = 0; b < 10 && rc == CGLI_OK; b++) {
        const int fid = (b < 5 ? CGLI_U : CGLI_UQ) + b % 5;
        const off_t off = (off_t)sizeof(double) * ((off_t)b * cf.nx_global * plane + (off_t)cf.gstart * plane);
        double *d = static_cast<double *>(dbuf);
        if (write) {
            launch_pack_q(c->g, c->f[fid], c->f[CGLI_J], d, c->stream);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(hbuf, d, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess)
                rc = CGLI_E_HIP;
            else if (!file_xfer(fd, hbuf, bytes, off, true))
                rc = CGLI_E_IO;
        } else {
            if (!file_xfer(fd, hbuf, bytes, off, false)) {
                rc = CGLI_E_IO;
                break;
            }
            if (hipMemcpyAsync(d, hbuf, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
                rc = CGLI_E_HIP;
                break;
            }
            launch_unpack_q(c->g, d, c->f[CGLI_J], c->f[fid], b < 5 ? c->f[CGLI_U_ + b] : nullptr, c->stream);
            // the host buffer is refilled by the next block: wait for this one
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = CGLI_E_HIP;
        }
    }
    if (fd >= 0 && close(fd) != 0 && rc == CGLI_OK) rc = CGLI_E_IO;
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == CGLI_OK) rc = CGLI_E_HIP;
    hipHostFree(hbuf);
    hipFree(dbuf);
    if (rc == CGLI_OK && !write) {
        mark(c, CGLI_U_, 15);
        ctx_channel(c).verify = true;
    }
    return rc;
}

extern "C" {

int cgli_write_restart(cgli_ctx *c, const char *path) { return restart_io(c, path, true); }
int cgli_read_restart(cgli_ctx *c, const char *path) { return restart_io(c, path, false); }

int cgli_timing_enable(cgli_ctx *c, int on)
{
    if (!c) return CGLI_E_ARG;
    c->timing = on != 0;
    return CGLI_OK;
}

int cgli_timing_select(cgli_ctx *c, unsigned mask)
{
    if (!c) return CGLI_E_ARG;
    c->tmask = mask;
    return CGLI_OK;
}

int cgli_timing_reset(cgli_ctx *c)
{
    if (!c) return CGLI_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < T_NCLASS; k++) c->ev_used[k] = 0;
    return CGLI_OK;
}

int cgli_timing_get(cgli_ctx *c, int cls, double *total_ms, long *count)
{
    if (!c || cls < 0 || cls >= T_NCLASS || !total_ms || !count) return CGLI_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    double t = 0;
    for (size_t q = 0; q < c->ev_used[cls]; q++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[cls][q].a, c->ev[cls][q].b));
        t += ms;
    }
    *total_ms = t;
    *count = (long)c->ev_used[cls];
    return CGLI_OK;
}

}  // extern "C"

namespace cgli {

// without the sponge term (the channel has none)
int ctx_rhs_sumsq(cgli_ctx *c, double deltaT, double e, double out[5])
{
    HIPCHK(hipSetDevice(c->device));
    int r = implicit_inputs(c);
    if (r) return r;
    const Geo &g = c->g;
    const long ncell = (long)g.nx * g.ny * g.nz;
    const long nb = (ncell + 255) / 256;
    r = ensure_scratch(c, 5 * nb + (long)g.NYY * g.NZZ + 64);
    if (!r) r = ensure_scratch(c, 5L * g.nx * g.ny + 16);
    if (r) return r;
    const hipStream_t s = c->stream;
    launch_rhs_diag(g, c->p, deltaT, e, 0, c5(c, CGLI_U_), c5(c, CGLI_U), c5(c, CGLI_UQ),
                    c->f[CGLI_J], m9(c, CGLI_MET), c5(c, CGLI_INFX), c5(c, CGLI_INFY), c5(c, CGLI_INFZ),
                    C4{{c->f[CGLI_VF], c->f[CGLI_VF + 1], c->f[CGLI_VF + 2], c->f[CGLI_VF + 3]}},
                    c5(c, CGLI_RES), f5(c, CGLI_MR), f5(c, CGLI_NL), f5(c, CGLI_NR), s, c->lusgs_mode < 3);
    double *dev = c->scratch + c->scratch_len - 8;
    launch_sumsq5(g, c5(c, CGLI_MR), c->scratch, dev, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->er_host, dev, sizeof(double) * 5, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int q = 0; q < 5; q++) {
        out[q] = c->er_host[q];
        if (!std::isfinite(out[q])) return CGLI_E_NONFINITE;
    }
    return CGLI_OK;
}

}  // namespace cgli
