// api.hip -- C-ABI implementation: contexts, device workspaces, the
// construction pipeline (znorm -> MFMA tiles -> FP64 recheck -> CSR) and the
// graph maps.  Declarations and reference citations are in include/bna.h.
#include "common.hpp"
#include "bna.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace bna;

namespace bna {
// Error band: a bound on |r_mfma - r64| for any pair, see band_for().
double band_for(int prec, int L_pad);
}  // namespace bna

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// ------------------------------------------------ no-allocation scratch
// bna_eigcent_noalloc keeps the power iteration's working memory here instead
// of in the context's workspace: two iterate vectors of BNA_NOALLOC_EC_MAX_N
// doubles, the per-workgroup partial sums and the iteration state.  Each GPU
// the library runs on holds its own copy (about 16.4 MiB) until the process
// ends.  In host code `scratch` names a host-side stand-in: the device address
// comes from hipGetSymbolAddress, once per device (scratch_setup), and only
// that address is handed to kernels.
constexpr size_t NA_X_BYTES = sizeof(double) * (size_t)BNA_NOALLOC_EC_MAX_N;
constexpr size_t NA_PART_BYTES = sizeof(double) * 3 * (size_t)EC_NPARTS_MAX;
constexpr size_t NA_SCRATCH_BYTES = 2 * NA_X_BYTES + NA_PART_BYTES + EC_STATE_MAX_BYTES;
static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex na_setup_mu;
static std::mutex na_call_mu[NA_MAX_DEVICES];  // one call at a time per GPU: they share the array
static bool na_looked_up[NA_MAX_DEVICES];
static void *na_ptr[NA_MAX_DEVICES];
static hipError_t na_err[NA_MAX_DEVICES];

// Resolves the device address of `scratch` on `device`, which must be the
// current device; later calls for the same device do nothing.
static void scratch_setup(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(na_setup_mu);
    if (na_looked_up[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    na_err[device] = e;
    na_ptr[device] = (e == hipSuccess) ? p : nullptr;
    na_looked_up[device] = true;
}

struct bna_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;
    size_t held = 0;  // workspace + series + csr bytes allocated through the context
    std::string err;
    bna_timings tm{};
    hipEvent_t ev[13] = {};  // [0..7], [12] builds, [8..9] graph maps, [10..11] bna_series_normalize
    bool znorm_pending = false;  // ev[10..11] bracket a normalisation not yet read into tm
    std::map<std::string, DevBuf> ws;
    int grid = 256;  // persistent correlation workgroups: CUs rounded up to a multiple of 8
    int ncu = 256;   // compute units of the device
    int grid_limit = 0;  // > 0: launches use at most this many (a multiple of 8)
    bool graphs = true;  // replay captured HIP graphs where a path has them (BNA_GRAPHS=0: off)
    // density bracket from a sample of tiles for subjects of at least this many
    // pairs (BNA_BRACKET_SAMPLE=0: always the full histogram pass;
    // BNA_BRACKET_SAMPLE_MIN=<pairs> changes the size; BNA_BRACKET_SAMPLE_SHIFT=<bins>
    // moves the sampled bracket up, which tests use to take the fallback)
    bool bracket_sample = true;
    int64_t bracket_sample_min = (int64_t)1 << 26;
    int bracket_sample_shift = 0;
    int64_t bracket_fallbacks = 0;  // sampled brackets that missed r* (the build redid the full pass)
    // tile traversal band height per (precision, L bucket), from the autotuner
    int band[3][4] = {{8, 8, 8, 8}, {8, 8, 8, 8}, {8, 8, 8, 8}};
    // tile table cache (single-subject builds): key of the cached table
    int tt_nb = -1, tt_nk = -1, tt_band = -1;
    const void *tt_panel = nullptr;
    int64_t tt_kstride = 0;
    int64_t tt_count = 0;
    // the context's own memory pool: allocations never wait for another
    // context's stream to reach a free
    hipMemPool_t pool = nullptr;
    bool own_pool = false;
    // chunked launches (serving): long kernels are issued in pieces of about
    // chunk_ms, each preceded by gate() (which may block)
    double chunk_ms = 0.0;
    void (*gate)(void *) = nullptr;
    void *gate_arg = nullptr;
    hipEvent_t chev[2] = {};
    hipEvent_t ecev[2] = {};  // the power iteration's two in-flight batches
    // pinned double buffer for host <-> device copies of pageable memory
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stev[2] = {};
    // pinned buffers of the build path, so none of its copies goes through the
    // runtime's staging of pageable memory (shared by every thread of the
    // process): [0] subject descriptors and [1] batch tile tables on the way in,
    // each guarded by the event of its last copy; [2] counters and histograms on
    // the way out, used between two stream synchronisations
    void *pin[3] = {nullptr, nullptr, nullptr};
    size_t pin_bytes[3] = {0, 0, 0};
    hipEvent_t pin_ev[2] = {};
    // batched builds: the emission launch and every subject's tails captured
    // as one HIP graph, replayed while the launch parameters stay the same
    std::string bg_key;
    hipGraph_t bg_graph = nullptr;
    hipGraphExec_t bg_exec = nullptr;
    int64_t bg_replays = 0, bg_captures = 0;
};

struct bna_series {
    int N = 0, L = 0, N_pad = 0, L_pad = 0, layout = 0;
    float *x = nullptr;
    bool own_x = false;
    float *xv = nullptr;  // time-major input only: the series voxel-major, row stride L_pad
    double *mu = nullptr, *sd = nullptr;
    uint8_t *valid = nullptr;
    unsigned char *panel = nullptr;  // k-blocked MFMA operands of prec_ready
    size_t panel_bytes = 0;
    int prec_ready = -1;  // precision whose operand panel is present
    size_t bytes = 0;
};

static thread_local std::string g_err;

static int fail(bna_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    g_err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), BNA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

static int dev_alloc(bna_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (ctx->budget && ctx->held + bytes > ctx->budget)
        return fail(ctx, BNA_ERR_DEVICE_NOMEM, "HBM budget exceeded: need " + std::to_string(bytes) +
                                                   " more bytes, holding " + std::to_string(ctx->held));
    // stream-ordered: no device-wide synchronisation, so contexts on other
    // streams keep running while this one allocates and frees
    hipError_t e = hipMallocFromPoolAsync(p, bytes, ctx->pool, ctx->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(ctx, BNA_ERR_DEVICE_NOMEM, std::string("hipMallocFromPoolAsync: ") + hipGetErrorString(e));
    }
    ctx->held += bytes;
    return BNA_OK;
}

// Output arrays belong to the returned bna_csr, not to the context.
static int out_alloc(bna_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    hipError_t e = hipMallocFromPoolAsync(p, bytes ? bytes : 256, ctx->pool, ctx->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(ctx, BNA_ERR_DEVICE_NOMEM, std::string("hipMallocFromPoolAsync (output): ") + hipGetErrorString(e));
    }
    return BNA_OK;
}

static void dev_free(bna_ctx *ctx, void *p, size_t bytes)
{
    if (!p) return;
    (void)hipFreeAsync(p, ctx->stream);
    if (bytes == 0) bytes = 256;
    ctx->held -= std::min(ctx->held, bytes);
}

// Grow-only named workspace buffer.
static int ws_get(bna_ctx *ctx, const char *name, size_t bytes, void **out)
{
    DevBuf &b = ctx->ws[name];
    if (b.p && b.bytes >= bytes) {
        *out = b.p;
        return BNA_OK;
    }
    if (b.p) {
        dev_free(ctx, b.p, b.bytes);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)256);
    int rc = dev_alloc(ctx, &b.p, want);
    if (rc) return rc;
    b.bytes = want;
    *out = b.p;
    return BNA_OK;
}

template <typename T>
static int ws_get_t(bna_ctx *ctx, const char *name, size_t count, T **out)
{
    void *p = nullptr;
    int rc = ws_get(ctx, name, count * sizeof(T), &p);
    *out = (T *)p;
    return rc;
}

static inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Copies between pageable host memory and the device go through the context's
// own pinned double buffer in 32 MB pieces (a host memcpy of one piece overlaps
// the DMA of the other), so they neither take the runtime's shared staging
// path nor wait for copies that other contexts have in flight.
static constexpr size_t STAGE_BYTES = (size_t)32 << 20;

static int stage_ready(bna_ctx *ctx)
{
    if (ctx->stage[0]) return BNA_OK;
    for (int b = 0; b < 2; b++) {
        HIPCHK(ctx, hipHostMalloc(&ctx->stage[b], STAGE_BYTES, hipHostMallocDefault));
        HIPCHK(ctx, hipEventCreateWithFlags(&ctx->stev[b], hipEventDisableTiming));
    }
    return BNA_OK;
}

// Grows pinned buffer `slot` to at least `bytes` (the caller has made sure no
// copy from or to it is in flight).
static int pin_grow(bna_ctx *ctx, int slot, size_t bytes)
{
    if (ctx->pin_bytes[slot] >= bytes) return BNA_OK;
    if (ctx->pin[slot]) {
        HIPCHK(ctx, hipHostFree(ctx->pin[slot]));
        ctx->pin[slot] = nullptr;
        ctx->pin_bytes[slot] = 0;
    }
    const size_t want = std::max(bytes, (size_t)64 << 10);
    HIPCHK(ctx, hipHostMalloc(&ctx->pin[slot], want, hipHostMallocDefault));
    ctx->pin_bytes[slot] = want;
    return BNA_OK;
}

// Host -> device copy through pinned buffer `slot` (0 or 1), queued on the
// context's stream; the buffer is reused only after its previous copy is done.
static int h2d_pinned(bna_ctx *ctx, int slot, void *dst, const void *src, size_t bytes)
{
    if (!ctx->pin_ev[slot]) HIPCHK(ctx, hipEventCreateWithFlags(&ctx->pin_ev[slot], hipEventDisableTiming));
    else HIPCHK(ctx, hipEventSynchronize(ctx->pin_ev[slot]));
    int rc = pin_grow(ctx, slot, bytes);
    if (rc) return rc;
    memcpy(ctx->pin[slot], src, bytes);
    HIPCHK(ctx, hipMemcpyAsync(dst, ctx->pin[slot], bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->pin_ev[slot], ctx->stream));
    return BNA_OK;
}

static int copy_h2d(bna_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    int rc = stage_ready(ctx);
    if (rc) return rc;
    if (bytes < ((size_t)1 << 20)) {  // one piece through the pinned stage
        memcpy(ctx->stage[0], src, bytes);
        HIPCHK(ctx, hipMemcpyAsync(dst, ctx->stage[0], bytes, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return BNA_OK;
    }
    size_t k = 0;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES, k++) {
        const int b = (int)(k & 1);
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        if (k >= 2) HIPCHK(ctx, hipEventSynchronize(ctx->stev[b]));  // its previous DMA is done
        // a chunked (long) build lets the copies of short ones go first
        if (ctx->chunk_ms > 0.0 && ctx->gate) ctx->gate(ctx->gate_arg);
        memcpy(ctx->stage[b], (const char *)src + off, n);
        HIPCHK(ctx, hipMemcpyAsync((char *)dst + off, ctx->stage[b], n, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->stev[b], ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BNA_OK;
}

static int copy_d2h(bna_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    int rc = stage_ready(ctx);
    if (rc) return rc;
    if (bytes < ((size_t)1 << 20)) {  // one piece through the pinned stage
        HIPCHK(ctx, hipMemcpyAsync(ctx->stage[0], src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        memcpy(dst, ctx->stage[0], bytes);
        return BNA_OK;
    }
    const size_t nch = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    auto issue = [&](size_t k) -> hipError_t {
        const size_t off = k * STAGE_BYTES, n = std::min(STAGE_BYTES, bytes - off);
        hipError_t e = hipMemcpyAsync(ctx->stage[k & 1], (const char *)src + off, n, hipMemcpyDeviceToHost,
                                      ctx->stream);
        return e == hipSuccess ? hipEventRecord(ctx->stev[k & 1], ctx->stream) : e;
    };
    HIPCHK(ctx, issue(0));
    for (size_t k = 0; k < nch; k++) {
        if (k + 1 < nch) HIPCHK(ctx, issue(k + 1));  // its buffer was emptied in the previous step
        HIPCHK(ctx, hipEventSynchronize(ctx->stev[k & 1]));
        const size_t off = k * STAGE_BYTES, n = std::min(STAGE_BYTES, bytes - off);
        memcpy((char *)dst + off, ctx->stage[k & 1], n);
    }
    return BNA_OK;
}

// Issues launch(first, count) over [0, n) in pieces of `per` units.  With a
// gate, a piece is issued only once the previous one has finished and the gate
// has let it through, so work that the gate waits for shares the device with
// at most the remainder of one piece; without one, two pieces stay in flight.
template <typename F>
static int run_chunks(bna_ctx *ctx, int64_t n, int64_t per, F launch)
{
    if (per <= 0 || per >= n || ctx->chunk_ms <= 0.0) {
        HIPCHK(ctx, launch((int64_t)0, n));
        return BNA_OK;
    }
    int k = 0;
    for (int64_t c0 = 0; c0 < n; c0 += per, k++) {
        if (ctx->gate && k >= 1) HIPCHK(ctx, hipEventSynchronize(ctx->chev[(k - 1) & 1]));
        if (!ctx->gate && k >= 2) HIPCHK(ctx, hipEventSynchronize(ctx->chev[k & 1]));
        if (ctx->gate) ctx->gate(ctx->gate_arg);
        HIPCHK(ctx, launch(c0, std::min(per, n - c0)));
        HIPCHK(ctx, hipEventRecord(ctx->chev[k & 1], ctx->stream));
    }
    return BNA_OK;
}

// Throughput figures that size the pieces: tile kernel ~3e6 tile-points per
// ms (tiles x L_pad), value refinement ~4e9 bytes per ms.
static constexpr double TILE_POINTS_PER_MS = 3.0e6;
static constexpr double REFINE_BYTES_PER_MS = 4.0e9;

// ------------------------------------------------------------------ context
extern "C" {

void bna_opts_default(bna_opts *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->thr_type = BNA_THR_R;
    o->thr = 0.0;
    o->weighted = 1;
    o->precision = BNA_BF16X3;
    o->recheck = 1;
    o->refine = 1;
}

const char *bna_status_string(int s)
{
    switch (s) {
    case BNA_OK: return "ok";
    case BNA_ERR_ARG: return "invalid argument";
    case BNA_ERR_HIP: return "HIP runtime error";
    case BNA_ERR_NOMEM: return "host out of memory";
    case BNA_ERR_DEVICE_NOMEM: return "device out of memory or over the HBM budget";
    case BNA_ERR_IO: return "file I/O error";
    case BNA_ERR_FORMAT: return "malformed file";
    case BNA_ERR_OVERFLOW: return "result exceeds the int32 CSR format";
    case BNA_ERR_NOT_FOUND: return "unknown request";
    case BNA_ERR_BUSY: return "request not finished";
    case BNA_ERR_CANCELLED: return "request cancelled";
    case BNA_ERR_DENSITY: return "density target not reachable";
    default: return "unknown status";
    }
}

const char *bna_last_error(bna_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int bna_create(bna_ctx **out, int device, size_t hbm_budget)
{
    if (!out) return fail(nullptr, BNA_ERR_ARG, "bna_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(nullptr, BNA_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(nullptr, BNA_ERR_ARG, "bna_create: bad device ordinal");
    bna_ctx *ctx = new (std::nothrow) bna_ctx();
    if (!ctx) return fail(nullptr, BNA_ERR_NOMEM, "bna_create: host allocation failed");
    ctx->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
        std::string m = std::string("bna_create: ") + hipGetErrorString(e);
        delete ctx;
        return fail(nullptr, BNA_ERR_HIP, m);
    }
    scratch_setup(device);  // a failure is reported by bna_eigcent_noalloc
    for (int b = 0; b < 4; b++) {
        hipEvent_t &ev = b < 2 ? ctx->chev[b] : ctx->ecev[b - 2];
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            std::string m = std::string("bna_create: ") + hipGetErrorString(e);
            bna_destroy(ctx);
            return fail(nullptr, BNA_ERR_HIP, m);
        }
    }
    for (auto &ev : ctx->ev) {
        if ((e = hipEventCreate(&ev)) != hipSuccess) {
            std::string m = std::string("bna_create: ") + hipGetErrorString(e);
            bna_destroy(ctx);
            return fail(nullptr, BNA_ERR_HIP, m);
        }
    }
    // a pool of the context's own: memory stays there between builds instead of
    // going back to the driver at every synchronisation, and a new allocation
    // never picks up memory whose free is still queued behind another stream
    hipMemPoolProps props;
    memset(&props, 0, sizeof(props));
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    if (hipMemPoolCreate(&ctx->pool, &props) == hipSuccess) {
        ctx->own_pool = true;
    } else {
        (void)hipGetLastError();
        if ((e = hipDeviceGetDefaultMemPool(&ctx->pool, device)) != hipSuccess) {
            std::string m = std::string("bna_create: no memory pool: ") + hipGetErrorString(e);
            bna_destroy(ctx);
            return fail(nullptr, BNA_ERR_HIP, m);
        }
    }
    {
        uint64_t keep = UINT64_MAX;
        int no = 0;
        (void)hipMemPoolSetAttribute(ctx->pool, hipMemPoolAttrReleaseThreshold, &keep);
        (void)hipMemPoolSetAttribute(ctx->pool, hipMemPoolReuseAllowInternalDependencies, &no);
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0) {
        ctx->grid = (ncu + 7) / 8 * 8;
        ctx->ncu = ncu;
    }
    {
        const char *g = std::getenv("BNA_GRAPHS");
        ctx->graphs = !(g && g[0] == '0');
        const char *bs = std::getenv("BNA_BRACKET_SAMPLE");
        ctx->bracket_sample = !(bs && bs[0] == '0');
        if (const char *bm = std::getenv("BNA_BRACKET_SAMPLE_MIN")) ctx->bracket_sample_min = std::atoll(bm);
        if (const char *bh = std::getenv("BNA_BRACKET_SAMPLE_SHIFT")) ctx->bracket_sample_shift = std::atoi(bh);
    }
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) hbm_budget = (size_t)(0.9 * (double)fr);
    }
    ctx->budget = hbm_budget;
    *out = ctx;
    return BNA_OK;
}

int bna_destroy(bna_ctx *ctx)
{
    if (!ctx) return BNA_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream) {
        for (auto &kv : ctx->ws)
            if (kv.second.p) (void)hipFreeAsync(kv.second.p, ctx->stream);
        (void)hipStreamSynchronize(ctx->stream);
    }
    // outputs still held by callers keep a destroyed pool alive until they are freed
    if (ctx->own_pool && ctx->pool) (void)hipMemPoolDestroy(ctx->pool);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (int b = 0; b < 2; b++) {
        if (ctx->stage[b]) (void)hipHostFree(ctx->stage[b]);
        if (ctx->stev[b]) (void)hipEventDestroy(ctx->stev[b]);
        if (ctx->chev[b]) (void)hipEventDestroy(ctx->chev[b]);
        if (ctx->ecev[b]) (void)hipEventDestroy(ctx->ecev[b]);
        if (ctx->pin_ev[b]) (void)hipEventDestroy(ctx->pin_ev[b]);
    }
    for (int b = 0; b < 3; b++)
        if (ctx->pin[b]) (void)hipHostFree(ctx->pin[b]);
    if (ctx->bg_exec) (void)hipGraphExecDestroy(ctx->bg_exec);
    if (ctx->bg_graph) (void)hipGraphDestroy(ctx->bg_graph);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BNA_OK;
}

int bna_synchronize(bna_ctx *ctx)
{
    if (!ctx) return fail(nullptr, BNA_ERR_ARG, "bna_synchronize: ctx is NULL");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BNA_OK;
}

int bna_get_timings(bna_ctx *ctx, bna_timings *t)
{
    if (!ctx || !t) return fail(ctx, BNA_ERR_ARG, "bna_get_timings: NULL argument");
    if (ctx->znorm_pending) {
        HIPCHK(ctx, hipEventSynchronize(ctx->ev[11]));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[10], ctx->ev[11]));
        ctx->tm.znorm_ms = ms;
        ctx->znorm_pending = false;
    }
    *t = ctx->tm;
    return BNA_OK;
}

int64_t bna_workspace_bytes(bna_ctx *ctx) { return ctx ? (int64_t)ctx->held : -1; }

int bna_set_chunking(bna_ctx *ctx, double chunk_ms, void (*gate)(void *), void *gate_arg)
{
    if (!ctx || chunk_ms < 0.0) return fail(ctx, BNA_ERR_ARG, "bna_set_chunking: bad argument");
    ctx->chunk_ms = chunk_ms;
    ctx->gate = gate;
    ctx->gate_arg = gate_arg;
    return BNA_OK;
}

int bna_set_grid_limit(bna_ctx *ctx, int max_workgroups)
{
    if (!ctx || max_workgroups < 0) return fail(ctx, BNA_ERR_ARG, "bna_set_grid_limit: bad argument");
    ctx->grid_limit = max_workgroups == 0 ? 0 : std::max(8, max_workgroups / 8 * 8);
    return BNA_OK;
}

int bna_grid_size(bna_ctx *ctx) { return ctx ? ctx->grid : -1; }

int bna_set_graphs(bna_ctx *ctx, int on)
{
    if (!ctx) return fail(ctx, BNA_ERR_ARG, "bna_set_graphs: ctx is NULL");
    ctx->graphs = on != 0;
    return BNA_OK;
}

int bna_batch_graph_stats(bna_ctx *ctx, int64_t *captures, int64_t *replays)
{
    if (!ctx) return fail(ctx, BNA_ERR_ARG, "bna_batch_graph_stats: ctx is NULL");
    if (captures) *captures = ctx->bg_captures;
    if (replays) *replays = ctx->bg_replays;
    return BNA_OK;
}

int bna_reserve(bna_ctx *ctx, size_t bytes)
{
    if (!ctx) return fail(nullptr, BNA_ERR_ARG, "bna_reserve: ctx is NULL");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = stage_ready(ctx);
    if (rc) return rc;
    for (int b = 0; b < 2; b++) memset(ctx->stage[b], 0, STAGE_BYTES);  // first touch now, not in a request
    if (bytes == 0) return BNA_OK;
    void *p = nullptr;
    HIPCHK(ctx, hipMallocFromPoolAsync(&p, bytes, ctx->pool, ctx->stream));
    HIPCHK(ctx, hipFreeAsync(p, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BNA_OK;
}

int bna_memory_info(bna_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes)
{
    if (!ctx || !free_bytes || !total_bytes) return fail(ctx, BNA_ERR_ARG, "bna_memory_info: NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    HIPCHK(ctx, hipMemGetInfo(&fr, &tot));
    *free_bytes = (int64_t)fr;
    *total_bytes = (int64_t)tot;
    return BNA_OK;
}

// ------------------------------------------------------------------ series
int bna_series_create(bna_ctx *ctx, const float *x, int N, int L, int layout, int x_on_device, bna_series **out)
{
    if (!ctx || !x || !out) return fail(ctx, BNA_ERR_ARG, "bna_series_create: NULL argument");
    *out = nullptr;
    if (N < 2 || L < 2) return fail(ctx, BNA_ERR_ARG, "bna_series_create: need N >= 2 and L >= 2");
    if (layout != BNA_TIME_MAJOR && layout != BNA_VOXEL_MAJOR) return fail(ctx, BNA_ERR_ARG, "bad layout");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    bna_series *s = new (std::nothrow) bna_series();
    if (!s) return fail(ctx, BNA_ERR_NOMEM, "bna_series_create: host allocation failed");
    s->N = N;
    s->L = L;
    s->layout = layout;
    s->N_pad = round_up(N, TILE);
    s->L_pad = std::max(LPAD_MIN, round_up(L, KPAD));
    const size_t xb = sizeof(float) * (size_t)N * L;
    if (x_on_device) {
        s->x = (float *)x;
        s->own_x = false;
    } else {
        int rc = dev_alloc(ctx, (void **)&s->x, xb);
        if (rc) {
            delete s;
            return rc;
        }
        s->own_x = true;
        s->bytes += xb;
        if ((rc = copy_h2d(ctx, s->x, x, xb))) {
            bna_series_destroy(ctx, s);
            return rc;
        }
    }
    *out = s;
    return BNA_OK;
}

static void series_free_norm(bna_ctx *ctx, bna_series *s)
{
    dev_free(ctx, s->xv, (size_t)s->N * s->L_pad * 4);
    dev_free(ctx, s->mu, (size_t)s->N * 8);
    dev_free(ctx, s->sd, (size_t)s->N * 8);
    dev_free(ctx, s->valid, (size_t)s->N);
    dev_free(ctx, s->panel, s->panel_bytes);
    s->xv = nullptr;
    s->mu = s->sd = nullptr;
    s->valid = nullptr;
    s->panel = nullptr;
    s->panel_bytes = 0;
    s->prec_ready = -1;
}

int bna_series_destroy(bna_ctx *ctx, bna_series *s)
{
    if (!s) return BNA_OK;
    if (!ctx) return fail(nullptr, BNA_ERR_ARG, "bna_series_destroy: ctx is NULL");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    series_free_norm(ctx, s);
    if (s->own_x) dev_free(ctx, s->x, sizeof(float) * (size_t)s->N * s->L);
    delete s;
    return BNA_OK;
}

static int series_normalize(bna_ctx *ctx, bna_series *s, int prec)
{
    if (prec < BNA_FP32 || prec > BNA_BF16X3) return fail(ctx, BNA_ERR_ARG, "bad precision");
    int rc;
    if (!s->mu) {
        if (s->layout == BNA_TIME_MAJOR && (rc = dev_alloc(ctx, (void **)&s->xv, (size_t)s->N * s->L_pad * 4)))
            return rc;
        if ((rc = dev_alloc(ctx, (void **)&s->mu, (size_t)s->N * 8))) return rc;
        if ((rc = dev_alloc(ctx, (void **)&s->sd, (size_t)s->N * 8))) return rc;
        if ((rc = dev_alloc(ctx, (void **)&s->valid, (size_t)s->N))) return rc;
    }
    // keep only the operand panel of this precision
    const size_t pb = (size_t)(s->L_pad / 16) * s->N_pad * panel_unit(prec);
    if (s->panel && s->panel_bytes != pb) {
        dev_free(ctx, s->panel, s->panel_bytes);
        s->panel = nullptr;
        s->panel_bytes = 0;
    }
    if (!s->panel) {
        if ((rc = dev_alloc(ctx, (void **)&s->panel, pb))) return rc;
        s->panel_bytes = pb;
    }
    HIPCHK(ctx, launch_znorm(s->x, s->layout, s->N, s->L, s->N_pad, s->L_pad, s->mu, s->sd, s->xv, s->valid,
                             prec, s->panel, ctx->stream));
    s->prec_ready = prec;
    return BNA_OK;
}

int bna_series_normalize(bna_ctx *ctx, bna_series *s, int precision)
{
    if (!ctx || !s) return fail(ctx, BNA_ERR_ARG, "bna_series_normalize: NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // asynchronous: work queued behind it on the context's stream (a build) is
    // ordered after it; its time is read at the next bna_get_timings
    HIPCHK(ctx, hipEventRecord(ctx->ev[10], ctx->stream));
    int rc = series_normalize(ctx, s, precision);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[11], ctx->stream));
    ctx->znorm_pending = true;
    return BNA_OK;
}

int bna_series_download(bna_ctx *ctx, const bna_series *s, double *z64, double *mu, double *sd)
{
    if (!ctx || !s) return fail(ctx, BNA_ERR_ARG, "bna_series_download: NULL argument");
    if (!s->mu) return fail(ctx, BNA_ERR_ARG, "bna_series_download: series not normalised");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (z64) {
        // z is not kept resident; it is materialised from the series for the copy
        const size_t zb = (size_t)s->N * s->L * 8;
        double *zd = nullptr;
        int rc = dev_alloc(ctx, (void **)&zd, zb);
        if (rc) return rc;
        hipError_t e = launch_zmat(s->xv ? s->xv : s->x, s->xv ? s->L_pad : s->L, s->N, s->L, s->mu, s->sd, zd,
                                   ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(z64, zd, zb, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        dev_free(ctx, zd, zb);
        if (e != hipSuccess) return fail(ctx, BNA_ERR_HIP, std::string("bna_series_download: ") + hipGetErrorString(e));
    }
    if (mu) HIPCHK(ctx, hipMemcpyAsync(mu, s->mu, (size_t)s->N * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (sd) HIPCHK(ctx, hipMemcpyAsync(sd, s->sd, (size_t)s->N * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return BNA_OK;
}

}  // extern "C"

// ------------------------------------------------------------ tile table
// Series-length buckets of the tuned tile traversal: up to 192, 320 and 640
// padded points, and longer.
static int l_bucket(int L_pad) { return L_pad <= 192 ? 0 : L_pad <= 320 ? 1 : L_pad <= 640 ? 2 : 3; }
static constexpr int BAND_CANDIDATES[4] = {4, 8, 16, 32};

// Upper-triangle tiles in bands of 8 tile-rows: within a band, tile-columns
// ascend and the band's rows vary fastest, so a run of consecutive items shares
// a few A panels (the band) and a few B panels (neighbouring columns).
static void make_tile_table(int nb, int subj, const void *panel, int64_t kstride, int nk, int G,
                            std::vector<TileItem> &out)
{
    for (int I0 = 0; I0 < nb; I0 += G) {
        const int I1 = std::min(nb, I0 + G);
        for (int J = I0; J < nb; J++)
            for (int I = I0; I < I1 && I <= J; I++) out.push_back(TileItem{panel, kstride, subj, I, J, nk});
    }
}

static int get_tile_table(bna_ctx *ctx, int nb, const void *panel, int64_t kstride, int nk, int G, TileItem **dev,
                          int64_t *count)
{
    if (ctx->tt_nb == nb && ctx->tt_panel == panel && ctx->tt_kstride == kstride && ctx->tt_nk == nk &&
        ctx->tt_band == G) {
        *dev = (TileItem *)ctx->ws["tiles"].p;
        *count = ctx->tt_count;
        return BNA_OK;
    }
    ctx->tt_nb = -1;
    std::vector<TileItem> t;
    make_tile_table(nb, 0, panel, kstride, nk, G, t);
    TileItem *d = nullptr;
    int rc = ws_get_t(ctx, "tiles", t.size(), &d);
    if (rc) return rc;
    if ((rc = copy_h2d(ctx, d, t.data(), t.size() * sizeof(TileItem)))) return rc;
    ctx->tt_nb = nb;
    ctx->tt_panel = panel;
    ctx->tt_kstride = kstride;
    ctx->tt_nk = nk;
    ctx->tt_band = G;
    ctx->tt_count = (int64_t)t.size();
    *dev = d;
    *count = ctx->tt_count;
    return BNA_OK;
}

// ------------------------------------------------------------ thresholds
namespace bna {

// Bound on |r_mfma - r64| per pair.  z64 has unit norm, so sum_l |z_i z_j| <= 1.
// Representation error of the operands: BF16X3 drops lo*lo and the residual
// below lo (<= 3.1 * 2^-16), BF16 keeps only hi (<= 2.01 * 2^-8), FP32 rounds
// z to float (<= 2.01 * 2^-24).  Accumulation: every product is added to an FP32
// accumulator; with at most one ulp (2^-23 relative) lost per addition the
// recursive-summation bound gives gamma_n = n u / (1 - n u), n = number of
// products per pair (3 L_pad for BF16X3, L_pad for BF16, 2 L_pad for FP32,
// where an FP32 product may round too), times the sum of |terms| (<= 1.02).
// A 1.25 safety factor and 1e-30 for flushed subnormals are added on top.
double band_for(int prec, int L_pad)
{
    const double u = std::ldexp(1.0, -23);
    double rep, n;
    if (prec == BNA_BF16X3) {
        rep = 3.1 * std::ldexp(1.0, -16);
        n = 3.0 * L_pad;
    } else if (prec == BNA_BF16) {
        rep = 2.01 * std::ldexp(1.0, -8);
        n = (double)L_pad;
    } else {
        rep = 2.01 * std::ldexp(1.0, -24);
        n = 2.0 * L_pad;
    }
    const double gam = n * u / (1.0 - n * u);
    return 1.25 * (rep + 1.02 * gam) + 1e-30;
}

}  // namespace bna

static float f_below(double t)  // largest float < t (strictly below)
{
    float f = (float)t;
    while ((double)f >= t) f = std::nextafter(f, -INFINITY);
    return f;
}
static float f_floor(double t)  // largest float <= t
{
    float f = (float)t;
    while ((double)f > t) f = std::nextafter(f, -INFINITY);
    return f;
}
static float f_ceil_above(double t)  // a float >= t, one ulp above the nearest
{
    float f = (float)t;
    while ((double)f < t) f = std::nextafter(f, INFINITY);
    return std::nextafter(f, INFINITY);
}

// ------------------------------------------------------------ build
// Histogram bin edge b of the tile kernel's r histogram over [-1, 1].
static double hist_edge(int b) { return -1.0 + (double)b / (HIST_BINS / 2); }

// Density bracket.  H(b) = number of pairs whose MFMA r falls in bins >= b.  The
// kernel's bin of v is floor((v + 1) * HIST_BINS / 2) in float arithmetic, so
//   #(v >= edge(b + 1)) <= H(b) <= #(v >= edge(b - 1)).
// With |r_mfma - r64| <= band and b the largest bin with H(b) >= k, the k-th
// largest r64, r*, satisfies  s_lo + band <= r* < s_hi - band  for
//   s_lo = edge(b - 1) - 2 band,  s_hi = edge(b + 2) + 2 band:
// then every pair with MFMA r >= s_hi has r64 > r* (a sure edge), every pair with
// MFMA r < s_lo has r64 < r* (never an edge), and all ties at r* lie in between.
struct DensityBracket {
    double s_lo, s_hi;
    int64_t emit_max, cand_max, valid_pairs;
};
static DensityBracket density_bracket(const unsigned long long *hist, int64_t k, double band)
{
    std::vector<int64_t> H(HIST_BINS + 2, 0);
    for (int b = HIST_BINS - 1; b >= 0; b--) H[b] = H[b + 1] + (int64_t)hist[b];
    DensityBracket d{};
    d.valid_pairs = H[0];
    int b = 0;
    for (int t = HIST_BINS - 1; t >= 0; t--)
        if (H[t] >= k) {
            b = t;
            break;
        }
    d.s_lo = b - 1 <= 0 ? -INFINITY : hist_edge(b - 1) - 2.0 * band;
    d.s_hi = b + 2 >= HIST_BINS ? INFINITY : hist_edge(b + 2) + 2.0 * band;
    // #(v >= s_lo) <= H(b_lo) for edge(b_lo + 1) <= s_lo; #(v >= s_hi) >= H(b_hi) for edge(b_hi - 1) >= s_hi
    int b_lo = 0, b_hi = HIST_BINS;
    if (std::isfinite(d.s_lo)) b_lo = std::max(0, std::min(HIST_BINS - 1, (int)std::floor((d.s_lo + 1.0) * (HIST_BINS / 2)) - 1));
    if (std::isfinite(d.s_hi)) b_hi = std::max(0, std::min(HIST_BINS, (int)std::ceil((d.s_hi + 1.0) * (HIST_BINS / 2)) + 1));
    d.emit_max = H[b_lo];
    d.cand_max = H[b_lo] - H[b_hi];
    return d;
}

static double decode_desc_key(unsigned long long k)
{
    uint64_t u = ~(uint64_t)k;
    const uint64_t bits = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    double r;
    memcpy(&r, &bits, sizeof r);
    return r;
}

// ------------------------------------------------------------ batched build
// One subject's construction inside a batch.  Threshold mode (k < 0): edge iff
// r > thr_eff, with pairs within the error band of thr_eff rechecked in FP64.
// Density mode (k >= 0): exactly k edges, from a histogram pass, a bracketed
// emission pass, an FP64 recheck of the bracket and an exact radix select.
// The tiles of every subject in a batch go through one histogram launch and
// one emission launch; the per-subject tails follow on the same stream.
struct Job {
    bna_series *s = nullptr;
    bna_opts o{};
    double thr_eff = 0.0;
    int64_t k = -1;  // density: edges wanted; -1: threshold mode
    int rc = BNA_OK;
    bool done = false;
    double band = 0.0, band_used = 0.0;
    std::string pre;  // workspace name prefix
    SubjDesc h{};
    unsigned long long *ctr = nullptr;  // [0] entries, [1] candidates, [2] max err, [3] flips
    unsigned long long *hist = nullptr;
    SelState *sel = nullptr;
    double *r64c = nullptr;
    int64_t ecap = 0, ccap = 0;
    int64_t *rowlen = nullptr, *R64 = nullptr, *bsum = nullptr, *off_rm = nullptr;
    int32_t *orig_rm = nullptr, *live_rm = nullptr;
    int32_t *R32 = nullptr, *Cd = nullptr;  // outputs, owned by the job until handed out
    float *Vd = nullptr;
    unsigned long long hc[8] = {};
    SelState hs{};
    int64_t nnz = 0;
    bool refined = false;  // values refined on the lists before placement
    // density: the bracket [s_lo, s_hi) of the emission pass, and whether it
    // came from a sample of the tiles (then checked against r* afterwards)
    double s_lo = -INFINITY, s_hi = INFINITY;
    bool sampled = false;
};

template <typename T>
static int job_ws(bna_ctx *ctx, Job &j, const char *name, size_t count, T **out)
{
    const std::string n = j.pre + name;
    return ws_get_t(ctx, n.c_str(), count, out);
}

static int job_prepare(bna_ctx *ctx, Job &j)
{
    bna_series *s = j.s;
    const int N = s->N, nb = s->N_pad / TILE;
    const size_t nfrag = (size_t)N * (nb + 1);
    int rc;
    int32_t *frag_orig, *frag_live;
    int64_t *frag_off;
    if ((rc = job_ws(ctx, j, "frag_orig", nfrag, &frag_orig)) || (rc = job_ws(ctx, j, "frag_live", nfrag, &frag_live)) ||
        (rc = job_ws(ctx, j, "frag_off", nfrag, &frag_off)) || (rc = job_ws(ctx, j, "counters", 8, &j.ctr)) ||
        (rc = job_ws(ctx, j, "rowlen", (size_t)N, &j.rowlen)) || (rc = job_ws(ctx, j, "R64", (size_t)N + 1, &j.R64)) ||
        (rc = job_ws(ctx, j, "scan_bsum", (size_t)scan_blocks(N) + 1, &j.bsum)) ||
        (rc = job_ws(ctx, j, "frag_orig_rm", nfrag, &j.orig_rm)) ||
        (rc = job_ws(ctx, j, "frag_live_rm", nfrag, &j.live_rm)) ||
        (rc = job_ws(ctx, j, "frag_off_rm", nfrag, &j.off_rm)))
        return rc;
    if (j.k >= 0) {
        if ((rc = job_ws(ctx, j, "hist", HIST_BINS, &j.hist)) || (rc = job_ws(ctx, j, "selstate", 1, &j.sel)))
            return rc;
    }
    // capacities: start from what the workspace already holds
    auto held = [&](const char *n, size_t unit) { return (int64_t)(ctx->ws[j.pre + n].bytes / unit); };
    j.ecap = std::max<int64_t>(1 << 20, held("entries", sizeof(Entry)));
    j.ccap = std::max<int64_t>(1 << 16, held("cands", sizeof(Candidate)));
    j.band = band_for(j.o.precision, s->L_pad);
    j.band_used = j.o.recheck ? j.band : 0.0;
    SubjDesc &h = j.h;
    h = SubjDesc{};
    h.a0 = s->panel;
    h.a1 = nullptr;
    h.N = N;
    h.N_pad = s->N_pad;
    h.L_pad = s->L_pad;
    h.nb = nb;
    h.frag_orig = frag_orig;
    h.frag_live = frag_live;
    h.frag_off = frag_off;
    h.entry_count = j.ctr;
    h.cand_count = j.ctr + 1;
    h.hist = nullptr;
    if (j.k < 0) {
        if (j.o.recheck) {
            h.keep_thr = f_below(j.thr_eff - j.band);
            h.sure_thr = f_ceil_above(j.thr_eff + j.band);
        } else {
            h.keep_thr = f_floor(j.thr_eff);
            h.sure_thr = h.keep_thr;
        }
    }
    return out_alloc(ctx, (void **)&j.R32, sizeof(int32_t) * ((size_t)N + 1));
}

// Tile table of a set of jobs: each job's band-ordered upper triangle, tagged
// with its index in the SubjDesc array.
static int upload_batch(bna_ctx *ctx, const std::vector<Job *> &js, SubjDesc **dsd, TileItem **tiles,
                        int64_t *ntiles)
{
    std::vector<SubjDesc> hd;
    for (Job *j : js) hd.push_back(j->h);
    int rc;
    if ((rc = ws_get_t(ctx, "batch_subjdesc", hd.size(), dsd))) return rc;
    if ((rc = h2d_pinned(ctx, 0, *dsd, hd.data(), hd.size() * sizeof(SubjDesc)))) return rc;
    auto kstride = [](const Job *j) { return (int64_t)j->s->N_pad * panel_unit(j->o.precision); };
    auto band = [&](const Job *j) { return ctx->band[j->o.precision][l_bucket(j->s->L_pad)]; };
    if (js.size() == 1)  // cached
        return get_tile_table(ctx, js[0]->s->N_pad / TILE, js[0]->s->panel, kstride(js[0]), js[0]->s->L_pad / 16,
                              band(js[0]), tiles, ntiles);
    std::vector<TileItem> t;
    for (size_t q = 0; q < js.size(); q++)
        make_tile_table(js[q]->s->N_pad / TILE, (int)q, js[q]->s->panel, kstride(js[q]), js[q]->s->L_pad / 16,
                        band(js[q]), t);
    if ((rc = ws_get_t(ctx, "batch_tiles", t.size(), tiles))) return rc;
    if ((rc = h2d_pinned(ctx, 1, *tiles, t.data(), t.size() * sizeof(TileItem)))) return rc;
    *ntiles = (int64_t)t.size();
    return BNA_OK;
}

static bool tile_sampled(int I, int J);
static int sampled_tile_count(int nb);
static int64_t tile_pairs(int N, int I, int J);

// upload_batch for the histogram pass when some jobs take a sample of their
// tiles: sampled jobs contribute only their sampled tiles; spairs[q] is the
// number of voxel pairs in job q's tiles of the table.
static int upload_batch_sampled(bna_ctx *ctx, const std::vector<Job *> &js, SubjDesc **dsd, TileItem **tiles,
                                int64_t *ntiles, std::vector<int64_t> &spairs)
{
    std::vector<SubjDesc> hd;
    for (Job *j : js) hd.push_back(j->h);
    int rc;
    if ((rc = ws_get_t(ctx, "batch_subjdesc", hd.size(), dsd))) return rc;
    if ((rc = h2d_pinned(ctx, 0, *dsd, hd.data(), hd.size() * sizeof(SubjDesc)))) return rc;
    std::vector<TileItem> t, one;
    spairs.assign(js.size(), 0);
    for (size_t q = 0; q < js.size(); q++) {
        const Job *j = js[q];
        one.clear();
        make_tile_table(j->s->N_pad / TILE, (int)q, j->s->panel, (int64_t)j->s->N_pad * panel_unit(j->o.precision),
                        j->s->L_pad / 16, ctx->band[j->o.precision][l_bucket(j->s->L_pad)], one);
        for (const TileItem &it : one)
            if (!j->sampled || tile_sampled(it.I, it.J)) {
                t.push_back(it);
                spairs[q] += tile_pairs(j->s->N, it.I, it.J);
            }
    }
    if ((rc = ws_get_t(ctx, "batch_tiles_hist", t.size(), tiles))) return rc;
    if ((rc = h2d_pinned(ctx, 1, *tiles, t.data(), t.size() * sizeof(TileItem)))) return rc;
    *ntiles = (int64_t)t.size();
    return BNA_OK;
}

static int corr_chunks(bna_ctx *ctx, int prec, int hist, SubjDesc *dsd, TileItem *tiles, int64_t nt,
                       const std::vector<Job *> &js)
{
    int Lmax = 32;
    for (Job *j : js) Lmax = std::max(Lmax, j->s->L_pad);
    const int64_t per = std::max<int64_t>(512, (int64_t)(ctx->chunk_ms * TILE_POINTS_PER_MS / Lmax) / 8 * 8);
    unsigned int *claim = nullptr;
    uint64_t *spill = nullptr;
    int rc = ws_get_t(ctx, "tile_claim", 8, &claim);
    if (!rc && !hist) rc = ws_get_t(ctx, "tile_spill", (size_t)corr_spill_records(prec, ctx->grid), &spill);
    if (rc) return rc;
    return run_chunks(ctx, nt, per, [&](int64_t c0, int64_t n) {
        const int grid = ctx->grid_limit > 0 ? std::min(ctx->grid, ctx->grid_limit) : ctx->grid;
        return launch_corr(prec, hist, dsd, tiles + c0, n, claim, spill, grid, ctx->stream);
    });
}

static void put_key(std::string &key, const void *p, size_t n) { key.append((const char *)p, n); }

// Density bracket from a sample of the tiles.  Tile (I, J) is in the sample when
// a hash of (I, J) is 0 mod BRACKET_SAMPLE_EVERY, so the sample spreads over the
// whole triangle.  pairs_all / pairs_sampled: the voxel pairs (i < j < N) of all
// tiles and of the sampled ones.
constexpr int BRACKET_SAMPLE_EVERY = 8;
static bool tile_sampled(int I, int J)
{
    uint32_t h = (uint32_t)I * 0x9E3779B1u ^ (uint32_t)J * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 13;
    return h % BRACKET_SAMPLE_EVERY == 0;
}
// Tiles of an nb x nb upper triangle that are in the sample.
static int sampled_tile_count(int nb)
{
    int c = 0;
    for (int J = 0; J < nb; J++)
        for (int I = 0; I <= J; I++) c += tile_sampled(I, J);
    return c;
}
static int64_t tile_pairs(int N, int I, int J)
{
    const int64_t ri = std::max(0, std::min(TILE, N - I * TILE)), rj = std::max(0, std::min(TILE, N - J * TILE));
    return I == J ? ri * (ri - 1) / 2 : ri * rj;
}

// Sampled form of density_bracket: k_s = k scaled to the sample, and the
// bracket spans the bins whose cumulative sampled count lies within a margin of
// k_s (0.5 % plus 12 standard deviations of a binomial count), widened by the
// error band as in the full form.  shift moves it up by that many bins.
static DensityBracket density_bracket_sampled(const unsigned long long *hist, int64_t k, double band,
                                              double frac, int shift)
{
    std::vector<int64_t> H(HIST_BINS + 2, 0);
    for (int b = HIST_BINS - 1; b >= 0; b--) H[b] = H[b + 1] + (int64_t)hist[b];
    const double ks = (double)k * frac;
    const double margin = 0.005 * ks + 12.0 * std::sqrt(ks + 1.0);
    auto top_bin = [&](double count) {  // largest bin with H >= count
        for (int t = HIST_BINS - 1; t >= 0; t--)
            if ((double)H[t] >= count) return t;
        return 0;
    };
    int b_top = top_bin(std::max(1.0, ks - margin)), b_bot = top_bin(ks + margin);
    b_top = std::max(0, std::min(HIST_BINS - 1, b_top + shift));
    b_bot = std::max(0, std::min(b_top, b_bot + shift));
    DensityBracket d{};
    d.s_lo = b_bot - 1 <= 0 ? -INFINITY : hist_edge(b_bot - 1) - 2.0 * band;
    d.s_hi = b_top + 2 >= HIST_BINS ? INFINITY : hist_edge(b_top + 2) + 2.0 * band;
    int b_lo = 0, b_hi = HIST_BINS;
    if (std::isfinite(d.s_lo)) b_lo = std::max(0, std::min(HIST_BINS - 1, (int)std::floor((d.s_lo + 1.0) * (HIST_BINS / 2)) - 1));
    if (std::isfinite(d.s_hi)) b_hi = std::max(0, std::min(HIST_BINS, (int)std::ceil((d.s_hi + 1.0) * (HIST_BINS / 2)) + 1));
    // capacities scaled up from the sample, with room for its error
    d.emit_max = (int64_t)((double)H[b_lo] / frac * 1.25) + 4096;
    d.cand_max = (int64_t)((double)(H[b_lo] - H[b_hi]) / frac * 1.25) + 4096;
    d.valid_pairs = -1;  // unknown from a sample
    return d;
}

static int build_batch(bna_ctx *ctx, std::vector<Job> &jobs, bool timing_from_norm)
{
    const int prec = jobs[0].o.precision;
    for (Job &j : jobs)
        if (j.o.precision != prec) return fail(ctx, BNA_ERR_ARG, "a batch shares one precision");
    float ms_corr = 0.f, ms_re = 0.f, ms_sel = 0.f, ms_rows = 0.f;
    int64_t launches = 0;
    int rc;
    for (size_t q = 0; q < jobs.size(); q++) {
        Job &j = jobs[q];
        j.pre = jobs.size() == 1 ? std::string() : "b" + std::to_string(q) + ":";
        if (j.k == 0) {
            j.rc = fail(ctx, BNA_ERR_DENSITY, "density: the percentage gives no edge (k = 0)");
            continue;
        }
        if ((j.rc = job_prepare(ctx, j))) return j.rc;
    }

    // Density subjects bracket r* from a histogram of the MFMA r (pass 1), then
    // emit, recheck and select (pass 2).  A large subject histograms a sample of
    // its tiles instead of all of them; its bracket is then checked against the
    // selected r* (the selection is exact when s_lo + band <= r* < s_hi - band:
    // every dropped pair lies below r* and every sure pair above it), and a
    // subject whose sampled bracket missed runs both passes again in full.
    for (int round = 0; round < 2; round++) {
        // pass 1 (density subjects): histogram of the MFMA r over every valid pair
        // (or over a sample of the tiles)
        std::vector<Job *> dj;
        for (Job &j : jobs)
            if (j.k > 0 && j.rc == BNA_OK && !j.done) dj.push_back(&j);
        if (!dj.empty()) {
            bool any_sampled = false;
            for (Job *j : dj) {
                const int64_t N = j->s->N;
                j->sampled = round == 0 && ctx->bracket_sample && N * (N - 1) / 2 >= ctx->bracket_sample_min &&
                             sampled_tile_count(j->s->N_pad / TILE) >= 8;
                any_sampled = any_sampled || j->sampled;
                j->h.hist = j->hist;
                j->h.keep_thr = INFINITY;
                j->h.sure_thr = INFINITY;
                HIPCHK(ctx, launch_zero_u64(j->hist, HIST_BINS, ctx->stream));
            }
            SubjDesc *dsd;
            TileItem *tiles;
            int64_t nt;
            std::vector<int64_t> spairs;
            if ((rc = any_sampled ? upload_batch_sampled(ctx, dj, &dsd, &tiles, &nt, spairs)
                                  : upload_batch(ctx, dj, &dsd, &tiles, &nt)))
                return rc;
            HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
            if ((rc = corr_chunks(ctx, prec, 1, dsd, tiles, nt, dj))) return rc;
            launches++;
            HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
            std::vector<unsigned long long> hh(HIST_BINS * dj.size());
            if ((rc = pin_grow(ctx, 2, hh.size() * sizeof(unsigned long long)))) return rc;
            unsigned long long *hp = (unsigned long long *)ctx->pin[2];
            for (size_t q = 0; q < dj.size(); q++)
                HIPCHK(ctx, hipMemcpyAsync(hp + q * HIST_BINS, dj[q]->hist, HIST_BINS * sizeof(unsigned long long),
                                           hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            memcpy(hh.data(), hp, hh.size() * sizeof(unsigned long long));
            float ms = 0.f;
            HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
            ms_corr += ms;
            for (size_t q = 0; q < dj.size(); q++) {
                Job &j = *dj[q];
                j.h.hist = nullptr;
                const int64_t N = j.s->N;
                DensityBracket br;
                if (j.sampled && spairs[q] > 0)
                    br = density_bracket_sampled(hh.data() + q * HIST_BINS, j.k, j.band_used,
                                                 (double)spairs[q] / (double)(N * (N - 1) / 2), ctx->bracket_sample_shift);
                else {
                    j.sampled = false;
                    br = density_bracket(hh.data() + q * HIST_BINS, j.k, j.band_used);
                    if (j.k > br.valid_pairs) {
                        j.rc = fail(ctx, BNA_ERR_DENSITY, "density: " + std::to_string(j.k) + " edges requested, " +
                                                              std::to_string(br.valid_pairs) + " valid pairs");
                        continue;
                    }
                }
                j.s_lo = br.s_lo;
                j.s_hi = br.s_hi;
                j.h.keep_thr = std::isfinite(br.s_lo) ? f_below(br.s_lo) : -INFINITY;  // emit v >= s_lo
                j.h.sure_thr = std::isfinite(br.s_hi) ? f_below(br.s_hi) : INFINITY;   // candidate v < s_hi
                j.ecap = std::max<int64_t>(j.ecap, 2 * br.emit_max + 4096);
                j.ccap = std::max<int64_t>(j.ccap, br.cand_max + 4096);
            }
        }

        // pass 2: emission, recheck (and selection), row lengths and scan; then one
        // host read of every job's counters.  A job whose lists overflowed grows them
        // and runs again (the recheck and the selection skip slots past capacity).
        for (int attempt = 0;; attempt++) {
            std::vector<Job *> act;
            for (Job &j : jobs)
                if (j.rc == BNA_OK && !j.done) act.push_back(&j);
            if (act.empty()) break;
            if (attempt == 3) return fail(ctx, BNA_ERR_DEVICE_NOMEM, "edge lists keep overflowing");
            for (Job *j : act) {
                Entry *entries;
                Candidate *cands;
                uint8_t *mirror = nullptr;
                const bool refine = j->o.weighted && j->o.refine;
                if ((refine && (rc = job_ws(ctx, *j, "mirror", (size_t)j->ecap, &mirror))) ||
                    (rc = job_ws(ctx, *j, "entries", (size_t)j->ecap, &entries)) ||
                    (rc = job_ws(ctx, *j, "cands", (size_t)j->ccap, &cands)) ||
                    (j->k > 0 && (rc = job_ws(ctx, *j, "cand_r64", (size_t)j->ccap, &j->r64c))))
                    return rc;
                j->h.entries = entries;
                j->h.entry_cap = j->ecap;
                j->h.mirror = mirror;
                j->h.cands = cands;
                j->h.cand_cap = j->ccap;
            }
            SubjDesc *dsd;
            TileItem *tiles;
            int64_t nt;
            if ((rc = upload_batch(ctx, act, &dsd, &tiles, &nt))) return rc;
            // per job, one pinned readback record: selection state, counters, nnz
            constexpr size_t RB = (sizeof(SelState) + sizeof(act[0]->hc) + sizeof(int64_t) + 63) / 64 * 64;
            if ((rc = pin_grow(ctx, 2, RB * act.size()))) return rc;
            unsigned char *rb = (unsigned char *)ctx->pin[2];
            // The segment from the emission launch to the counter readbacks.  In a
            // batch it is captured as one HIP graph (stage events are not recorded
            // inside; the tile kernel's time then covers the tails too) and
            // replayed while its launch parameters, the key, stay the same.
            const bool graph = ctx->graphs && jobs.size() > 1 && ctx->chunk_ms <= 0.0;
            auto segment = [&](bool events) -> int {
                for (Job *j : act) HIPCHK(ctx, launch_zero_u64(j->ctr, 8, ctx->stream));
                if (events) HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
                int rc2;
                if ((rc2 = corr_chunks(ctx, prec, 0, dsd, tiles, nt, act))) return rc2;
                if (events) HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
                for (Job *j : act) {
                    bna_series *s = j->s;
                    const int N = s->N, nb = s->N_pad / TILE;
                    const float *xs = s->xv ? s->xv : s->x;
                    const int64_t ld = s->xv ? s->L_pad : s->L;
                    const int mode = j->k > 0 ? (j->o.recheck ? 1 : 2) : 0;
                    HIPCHK(ctx, launch_recheck(xs, ld, s->mu, s->sd, N, s->L, nb, j->h.cands, j->ctr + 1, j->h.cand_cap,
                                               j->thr_eff, j->h.frag_live, j->h.entries, j->h.entry_cap, j->ctr + 2, mode,
                                               j->r64c, ctx->stream));
                }
                if (events) HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
                for (size_t q = 0; q < act.size(); q++) {
                    Job *j = act[q];
                    if (j->k > 0) {
                        HIPCHK(ctx, launch_select(j->sel, j->ctr, j->h.cands, j->r64c, j->h.cand_cap, j->k, j->s->N,
                                                  j->h.frag_live, j->h.entries, j->h.entry_cap, ctx->stream));
                        HIPCHK(ctx, hipMemcpyAsync(rb + q * RB, j->sel, sizeof(SelState), hipMemcpyDeviceToHost,
                                                   ctx->stream));
                    }
                }
                if (events) HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
                for (size_t q = 0; q < act.size(); q++) {
                    Job *j = act[q];
                    const int N = j->s->N, nb = j->s->N_pad / TILE;
                    HIPCHK(ctx, launch_frag_rows(j->h.frag_orig, j->h.frag_live, j->h.frag_off, N, nb, j->orig_rm,
                                                 j->live_rm, j->off_rm, j->rowlen, ctx->stream));
                    HIPCHK(ctx, launch_scan(j->rowlen, N, j->R64, nullptr, j->bsum, ctx->stream));
                    HIPCHK(ctx, hipMemcpyAsync(rb + q * RB + sizeof(SelState), j->ctr, sizeof(j->hc),
                                               hipMemcpyDeviceToHost, ctx->stream));
                    HIPCHK(ctx, hipMemcpyAsync(rb + q * RB + sizeof(SelState) + sizeof(j->hc), j->R64 + N,
                                               sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
                }
                if (events) HIPCHK(ctx, hipEventRecord(ctx->ev[12], ctx->stream));  // the CSR's row pass is done
                return BNA_OK;
            };
            bool replayed = false;
            if (graph) {
                // the tile kernel's workspaces exist before the capture, so nothing
                // is allocated inside it
                unsigned int *claim_ws = nullptr;
                uint64_t *spill_ws = nullptr;
                if ((rc = ws_get_t(ctx, "tile_claim", 8, &claim_ws)) ||
                    (rc = ws_get_t(ctx, "tile_spill", (size_t)corr_spill_records(prec, ctx->grid), &spill_ws)))
                    return rc;
                // the key: every value the segment's launches take
                std::string key;
                put_key(key, &claim_ws, sizeof claim_ws);
                put_key(key, &spill_ws, sizeof spill_ws);
                auto put = [&](const void *p, size_t n) { put_key(key, p, n); };
                const int grid = ctx->grid_limit > 0 ? std::min(ctx->grid, ctx->grid_limit) : ctx->grid;
                put(&prec, sizeof prec);
                put(&grid, sizeof grid);
                put(&dsd, sizeof dsd);
                put(&tiles, sizeof tiles);
                put(&nt, sizeof nt);
                put(&rb, sizeof rb);
                for (Job *j : act) {
                    put(&j->h, sizeof(SubjDesc));
                    const void *ptrs[] = {j->ctr, j->sel, j->r64c, j->rowlen, j->R64, j->bsum, j->orig_rm,
                                          j->live_rm, j->off_rm, j->s->x, j->s->xv, j->s->mu, j->s->sd};
                    put(ptrs, sizeof ptrs);
                    const int ints[] = {j->s->N, j->s->L, j->s->L_pad, j->s->N_pad, j->o.recheck};
                    put(ints, sizeof ints);
                    put(&j->thr_eff, sizeof j->thr_eff);
                    put(&j->k, sizeof j->k);
                }
                if (ctx->bg_exec && key == ctx->bg_key) {
                    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
                    HIPCHK(ctx, hipGraphLaunch(ctx->bg_exec, ctx->stream));
                    ctx->bg_replays++;
                    replayed = true;
                } else {
                    if (ctx->bg_exec) (void)hipGraphExecDestroy(ctx->bg_exec);
                    if (ctx->bg_graph) (void)hipGraphDestroy(ctx->bg_graph);
                    ctx->bg_exec = nullptr;
                    ctx->bg_graph = nullptr;
                    ctx->bg_key.clear();
                    HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                    const int rc2 = segment(false);
                    hipGraph_t g = nullptr;
                    const hipError_t ee = hipStreamEndCapture(ctx->stream, &g);
                    if (rc2) {
                        if (g) (void)hipGraphDestroy(g);
                        return rc2;
                    }
                    HIPCHK(ctx, ee);
                    hipGraphExec_t ex = nullptr;
                    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                    if (ei != hipSuccess) {
                        (void)hipGraphDestroy(g);
                        HIPCHK(ctx, ei);
                    }
                    ctx->bg_graph = g;
                    ctx->bg_exec = ex;
                    ctx->bg_key = key;
                    ctx->bg_captures++;
                    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
                    HIPCHK(ctx, hipGraphLaunch(ex, ctx->stream));
                    replayed = true;
                }
                // the whole segment is timed as the tile kernel's
                HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
                HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
                HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
                HIPCHK(ctx, hipEventRecord(ctx->ev[12], ctx->stream));
            } else if ((rc = segment(true))) {
                return rc;
            }
            (void)replayed;
            launches++;
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            for (size_t q = 0; q < act.size(); q++) {
                Job *j = act[q];
                if (j->k > 0) memcpy(&j->hs, rb + q * RB, sizeof(SelState));
                memcpy(j->hc, rb + q * RB + sizeof(SelState), sizeof(j->hc));
                memcpy(&j->nnz, rb + q * RB + sizeof(SelState) + sizeof(j->hc), sizeof(int64_t));
            }
            float ms = 0.f;
            HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
            ms_corr += ms;
            HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]));
            ms_re += ms;
            HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
            ms_sel += ms;
            HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[12]));
            ms_rows += ms;
            for (Job *j : act) {
                const unsigned long long *hc = j->hc;
                const bool e_over = (int64_t)hc[0] > j->ecap, c_over = (int64_t)hc[1] > j->ccap;
                if (!e_over && !c_over) {
                    j->done = true;
                    continue;
                }
                if (e_over) j->ecap = (int64_t)(hc[0] + hc[0] / 8 + 1024);
                if (c_over) j->ccap = (int64_t)(hc[1] + hc[1] / 8 + 1024);
            }
        }

        // a sampled bracket must contain r* with the band to spare on both sides
        bool redo = false;
        for (Job &j : jobs) {
            if (j.rc != BNA_OK || j.k <= 0 || !j.sampled) continue;
            j.sampled = false;
            const double rs = decode_desc_key(j.hs.kmax);
            const bool ok = !j.hs.err && j.nnz == 2 * j.k && j.s_lo + j.band_used <= rs && rs < j.s_hi - j.band_used;
            if (!ok) {
                j.done = false;
                redo = true;
                ctx->bracket_fallbacks++;
            }
        }
        if (!redo) break;
    }

    // value refinement on the lists, then placement, of every finished job
    HIPCHK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    for (size_t q = 0; q < jobs.size(); q++) {
        Job &j = jobs[q];
        bna_series *s = j.s;
        if (j.rc != BNA_OK || !j.o.weighted || !j.o.refine || !j.h.mirror) continue;
        if (j.k > 0 && (j.hs.err || j.nnz != 2 * j.k)) continue;  // fails below
        const int64_t ld = s->xv ? s->L_pad : s->L;
        if (!refine_lists_supported(s->xv ? s->xv : s->x, ld, s->L)) continue;  // refined after placement
        j.refined = true;
        const double bytes = (double)j.nnz / 2.0 * s->L * 4.0 + (double)s->N * s->L * 4.0;
        const int64_t per = (int64_t)std::max(64.0, s->N * (ctx->chunk_ms * REFINE_BYTES_PER_MS / bytes));
        if ((rc = run_chunks(ctx, s->N, per, [&](int64_t r0, int64_t n) {
                 return launch_refine_lists(j.h.entries, j.h.mirror, j.orig_rm, j.off_rm, j.h.frag_off,
                                            s->xv ? s->xv : s->x, ld, s->mu, s->sd, s->N, s->L, s->N_pad / TILE,
                                            (int)r0, (int)(r0 + n), ctx->stream);
             })))
            return rc;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
    for (size_t q = 0; q < jobs.size(); q++) {
        Job &j = jobs[q];
        if (j.rc != BNA_OK) continue;
        if (j.k > 0 && (j.hs.err || j.nnz != 2 * j.k)) {
            j.rc = fail(ctx, BNA_ERR_DENSITY, "density selection inconsistent (err " + std::to_string(j.hs.err) +
                                                  ", nnz " + std::to_string(j.nnz) + ", k " + std::to_string(j.k) + ")");
            continue;
        }
        if (j.nnz > 0x7fffffffLL) {
            j.rc = fail(ctx, BNA_ERR_OVERFLOW, "nnz exceeds int32 (" + std::to_string(j.nnz) + ")");
            continue;
        }
        if ((j.rc = out_alloc(ctx, (void **)&j.Cd, sizeof(int32_t) * (size_t)std::max<int64_t>(j.nnz, 1))))
            continue;
        if (j.o.weighted &&
            (j.rc = out_alloc(ctx, (void **)&j.Vd, sizeof(float) * (size_t)std::max<int64_t>(j.nnz, 1))))
            continue;
        bna_series *s = j.s;
        HIPCHK(ctx, launch_narrow_rows(j.R64, (int64_t)s->N + 1, j.R32, ctx->stream));
        HIPCHK(ctx, launch_place(j.orig_rm, j.live_rm, j.off_rm, j.h.entries, j.R64, s->N, s->N_pad / TILE, j.Cd,
                                 j.Vd, ctx->stream));
    }
    // shapes the list form does not take: refinement of the placed CSR
    for (size_t q = 0; q < jobs.size(); q++) {
        Job &j = jobs[q];
        bna_series *s = j.s;
        if (j.rc == BNA_OK && j.Vd && j.o.refine && !j.refined && j.nnz > 0) {
            const double bytes = (double)j.nnz / 2.0 * s->L * 4.0 + (double)s->N * s->L * 4.0;
            const int64_t per = (int64_t)std::max(64.0, s->N * (ctx->chunk_ms * REFINE_BYTES_PER_MS / bytes));
            if ((rc = run_chunks(ctx, s->N, per, [&](int64_t r0, int64_t n) {
                     return launch_refine(j.R32, j.Cd, j.Vd, s->xv ? s->xv : s->x, s->xv ? s->L_pad : s->L, s->mu,
                                          s->sd, s->N, s->L, (int)r0, (int)(r0 + n), j.nnz, ctx->stream);
                 })))
                return rc;
        }
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[7], ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));

    float ms_csr = 0.f, ms_ref = 0.f, ms_tot = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms_ref, ctx->ev[5], ctx->ev[6]));
    HIPCHK(ctx, hipEventElapsedTime(&ms_csr, ctx->ev[6], ctx->ev[7]));
    HIPCHK(ctx, hipEventElapsedTime(&ms_tot, timing_from_norm ? ctx->ev[0] : ctx->ev[1], ctx->ev[7]));
    ctx->tm.corr_ms = ms_corr;
    ctx->tm.recheck_ms = ms_re;
    ctx->tm.select_ms = ms_sel;
    ctx->tm.csr_ms = ms_rows + ms_csr;
    ctx->tm.refine_ms = ms_ref;
    ctx->tm.total_ms = ms_tot;
    ctx->tm.corr_launches = launches;
    return BNA_OK;
}

// Hands a finished job's arrays to out; a failed job's arrays are freed.
static void job_result(bna_ctx *ctx, Job &j, bna_csr *out)
{
    bna_series *s = j.s;
    memset(out, 0, sizeof(*out));
    if (j.rc != BNA_OK) {
        if (j.R32) (void)hipFreeAsync(j.R32, ctx->stream);
        if (j.Cd) (void)hipFreeAsync(j.Cd, ctx->stream);
        if (j.Vd) (void)hipFreeAsync(j.Vd, ctx->stream);
        j.R32 = j.Cd = nullptr;
        j.Vd = nullptr;
        return;
    }
    out->N = s->N;
    out->nnz = j.nnz;
    out->R = j.R32;
    out->C = j.Cd;
    out->V = j.Vd;
    j.R32 = j.Cd = nullptr;
    j.Vd = nullptr;
    out->on_device = 1;
    out->thr_used = j.k > 0 ? decode_desc_key(j.hs.kmax) : j.thr_eff;
    const double M1 = (double)s->N * (s->N - 1) / 2.0;
    out->density_pct = 100.0 * (double)j.nnz / M1 / 2.0;
    out->n_candidates = (int64_t)j.hc[1];
    out->n_flips = j.k > 0 ? 0 : (int64_t)j.hc[3];
    out->band = j.band_used;
    double me = 0.0;
    memcpy(&me, &j.hc[2], sizeof(double));
    out->max_err = me;
}

extern "C" {

// Sets up a job from options; normalises the series when its operand panel is
// not of the requested precision.
static int job_init(bna_ctx *ctx, Job &j, bna_series *s, const bna_opts *o, bool *normed)
{
    if (o->precision < BNA_FP32 || o->precision > BNA_BF16X3) return fail(ctx, BNA_ERR_ARG, "bad precision");
    j.s = s;
    j.o = *o;
    if (o->thr_type == BNA_THR_R) {
        j.thr_eff = o->thr - (o->ref_compat_eps ? 1e-6 : 0.0);
        j.k = -1;
    } else if (o->thr_type == BNA_THR_DENSITY) {
        if (!(o->thr >= 0.0 && o->thr <= 100.0)) return fail(ctx, BNA_ERR_ARG, "density must be in [0, 100] percent");
        j.k = bna_density_to_k(s->N, (float)o->thr);
    } else {
        return fail(ctx, BNA_ERR_ARG, "bad thr_type");
    }
    if (s->prec_ready != o->precision) {
        int rc = series_normalize(ctx, s, o->precision);
        if (rc) return rc;
        *normed = true;
    }
    return BNA_OK;
}

int bna_build_csr_series(bna_ctx *ctx, bna_series *s, const bna_opts *o, bna_csr *out)
{
    if (!ctx || !s || !o || !out) return fail(ctx, BNA_ERR_ARG, "bna_build_csr_series: NULL argument");
    memset(out, 0, sizeof(*out));
    HIPCHK(ctx, hipSetDevice(ctx->device));
    memset(&ctx->tm, 0, sizeof(ctx->tm));
    bool normed = false;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    std::vector<Job> jobs(1);
    int rc = job_init(ctx, jobs[0], s, o, &normed);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    rc = build_batch(ctx, jobs, normed);
    if (rc == BNA_OK) rc = jobs[0].rc;
    job_result(ctx, jobs[0], out);
    if (rc) return rc;
    if (normed) {
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->tm.znorm_ms = ms;
        ctx->znorm_pending = false;
    }
    return BNA_OK;
}

int bna_build_csr_batch(bna_ctx *ctx, bna_series *const *series, const bna_opts *opts, int n, bna_csr *out,
                        int *status)
{
    if (!ctx || !series || !opts || !out || n < 1) return fail(ctx, BNA_ERR_ARG, "bna_build_csr_batch: bad argument");
    for (int q = 0; q < n; q++) {
        memset(&out[q], 0, sizeof(out[q]));
        if (status) status[q] = BNA_OK;
        if (!series[q]) return fail(ctx, BNA_ERR_ARG, "bna_build_csr_batch: NULL series");
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    memset(&ctx->tm, 0, sizeof(ctx->tm));
    bool normed = false;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    std::vector<Job> jobs(n);
    for (int q = 0; q < n; q++) {
        int rc = job_init(ctx, jobs[q], series[q], &opts[q], &normed);
        if (rc) return rc;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    int rc = build_batch(ctx, jobs, normed);
    int first = BNA_OK;
    for (int q = 0; q < n; q++) {
        if (rc && jobs[q].rc == BNA_OK) jobs[q].rc = rc;
        if (status) status[q] = jobs[q].rc;
        if (first == BNA_OK) first = jobs[q].rc;
        job_result(ctx, jobs[q], &out[q]);
    }
    if (normed && rc == BNA_OK) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->tm.znorm_ms = ms;
        ctx->znorm_pending = false;
    }
    return rc ? rc : first;
}

int bna_trim(bna_ctx *ctx)
{
    if (!ctx) return fail(nullptr, BNA_ERR_ARG, "bna_trim: ctx is NULL");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &kv : ctx->ws) dev_free(ctx, kv.second.p, kv.second.bytes);
    ctx->ws.clear();
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemPoolTrimTo(ctx->pool, 0));
    ctx->tt_nb = -1;
    ctx->tt_count = 0;
    return BNA_OK;
}

int bna_csr_release(bna_ctx *ctx, bna_csr *g)
{
    if (!ctx || !g) return fail(ctx, BNA_ERR_ARG, "bna_csr_release: NULL argument");
    if (!g->on_device) {
        bna_csr_free(g);
        return BNA_OK;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (g->R) HIPCHK(ctx, hipFreeAsync(g->R, ctx->stream));
    if (g->C) HIPCHK(ctx, hipFreeAsync(g->C, ctx->stream));
    if (g->V) HIPCHK(ctx, hipFreeAsync(g->V, ctx->stream));
    g->R = g->C = nullptr;
    g->V = nullptr;
    g->nnz = 0;
    return BNA_OK;
}

void bna_csr_free(bna_csr *g)
{
    if (!g) return;
    if (g->on_device) {
        // stream-ordered frees on the null stream: the library's streams are
        // non-blocking, so this waits for nothing they run; the caller must not
        // have work pending on these arrays
        if (g->R) (void)hipFreeAsync(g->R, nullptr);
        if (g->C) (void)hipFreeAsync(g->C, nullptr);
        if (g->V) (void)hipFreeAsync(g->V, nullptr);
    } else {
        free(g->R);
        free(g->C);
        free(g->V);
    }
    g->R = nullptr;
    g->C = nullptr;
    g->V = nullptr;
    g->nnz = 0;
}

int bna_csr_to_host(bna_ctx *ctx, const bna_csr *dev, bna_csr *out)
{
    if (!ctx || !dev || !out) return fail(ctx, BNA_ERR_ARG, "bna_csr_to_host: NULL argument");
    *out = *dev;
    out->on_device = 0;
    const size_t nz = (size_t)std::max<int64_t>(dev->nnz, 1);
    out->R = (int32_t *)malloc(sizeof(int32_t) * ((size_t)dev->N + 1));
    out->C = (int32_t *)malloc(sizeof(int32_t) * nz);
    out->V = dev->V ? (float *)malloc(sizeof(float) * nz) : nullptr;
    if (!out->R || !out->C || (dev->V && !out->V)) {
        free(out->R);
        free(out->C);
        free(out->V);
        out->R = out->C = nullptr;
        out->V = nullptr;
        return fail(ctx, BNA_ERR_NOMEM, "bna_csr_to_host: host allocation failed");
    }
    if (!dev->on_device) {
        memcpy(out->R, dev->R, sizeof(int32_t) * ((size_t)dev->N + 1));
        memcpy(out->C, dev->C, sizeof(int32_t) * (size_t)dev->nnz);
        if (dev->V) memcpy(out->V, dev->V, sizeof(float) * (size_t)dev->nnz);
        return BNA_OK;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int rc = copy_d2h(ctx, out->R, dev->R, sizeof(int32_t) * ((size_t)dev->N + 1));
    if (!rc && dev->nnz > 0) rc = copy_d2h(ctx, out->C, dev->C, sizeof(int32_t) * (size_t)dev->nnz);
    if (!rc && dev->nnz > 0 && dev->V) rc = copy_d2h(ctx, out->V, dev->V, sizeof(float) * (size_t)dev->nnz);
    if (rc) {
        free(out->R);
        free(out->C);
        free(out->V);
        out->R = out->C = nullptr;
        out->V = nullptr;
    }
    return rc;
}

int bna_build_csr(bna_ctx *ctx, const float *x, int N, int L, int layout, const bna_opts *o, bna_csr *out)
{
    if (!ctx || !x || !o || !out) return fail(ctx, BNA_ERR_ARG, "bna_build_csr: NULL argument");
    memset(out, 0, sizeof(*out));
    bna_series *s = nullptr;
    int rc = bna_series_create(ctx, x, N, L, layout, 0, &s);
    if (rc) return rc;
    bna_csr dev{};
    rc = bna_build_csr_series(ctx, s, o, &dev);
    if (rc == BNA_OK) {
        rc = bna_csr_to_host(ctx, &dev, out);
        bna_csr_release(ctx, &dev);
    }
    bna_series_destroy(ctx, s);
    return rc;
}

// ------------------------------------------------------------ graph maps
// Makes a device view of g (uploading host arrays into the workspace).
static int device_view(bna_ctx *ctx, const bna_csr *g, const int32_t **R, const int32_t **C, const float **V)
{
    if (g->on_device) {
        *R = g->R;
        *C = g->C;
        *V = g->V;
        return BNA_OK;
    }
    // host arrays are checked before any kernel indexes with them
    if (g->R[0] != 0 || (int64_t)g->R[g->N] != g->nnz) return fail(ctx, BNA_ERR_FORMAT, "CSR: R[0] != 0 or R[N] != nnz");
    for (int32_t k = 0; k < g->N; k++)
        if (g->R[k + 1] < g->R[k]) return fail(ctx, BNA_ERR_FORMAT, "CSR: R not monotone");
    for (int64_t k = 0; k < g->nnz; k++)
        if (g->C[k] < 0 || g->C[k] >= g->N) return fail(ctx, BNA_ERR_FORMAT, "CSR: column out of range");
    int32_t *r, *c;
    float *v = nullptr;
    int rc;
    const size_t nz = (size_t)std::max<int64_t>(g->nnz, 1);
    if ((rc = ws_get_t(ctx, "g_R", (size_t)g->N + 1, &r))) return rc;
    if ((rc = ws_get_t(ctx, "g_C", nz, &c))) return rc;
    if (g->V && (rc = ws_get_t(ctx, "g_V", nz, &v))) return rc;
    if ((rc = copy_h2d(ctx, r, g->R, sizeof(int32_t) * ((size_t)g->N + 1)))) return rc;
    if (g->nnz > 0) {
        if ((rc = copy_h2d(ctx, c, g->C, sizeof(int32_t) * (size_t)g->nnz))) return rc;
        if (g->V && (rc = copy_h2d(ctx, v, g->V, sizeof(float) * (size_t)g->nnz))) return rc;
    }
    *R = r;
    *C = c;
    *V = v;
    return BNA_OK;
}

int bna_degree(bna_ctx *ctx, const bna_csr *g, float *deg)
{
    if (!ctx || !g || !deg) return fail(ctx, BNA_ERR_ARG, "bna_degree: NULL argument");
    if (g->N < 1 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_degree: empty graph");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    float *d;
    if ((rc = ws_get_t(ctx, "deg", (size_t)g->N, &d))) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    HIPCHK(ctx, launch_degree(R, V, g->N, g->nnz, d, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    int rc2 = copy_d2h(ctx, deg, d, sizeof(float) * (size_t)g->N);  // through pinned memory; synchronises
    if (rc2) return rc2;
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.degree_ms = ms;
    return BNA_OK;
}

int bna_modularity(bna_ctx *ctx, const bna_csr *g, int32_t *labels, double *Q, int *n_modules)
{
    if (!ctx || !g || !labels || !Q) return fail(ctx, BNA_ERR_ARG, "bna_modularity: NULL argument");
    if (g->N < 1 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_modularity: empty graph");
    if (g->nnz < 2) return fail(ctx, BNA_ERR_ARG, "bna_modularity: the graph has no edge");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    const int N = g->N;
    const size_t nz = (size_t)g->nnz;
    NmBuffers b;
    if ((rc = ws_get_t(ctx, "nm_AD", (size_t)N, &b.AD))) return rc;
    if ((rc = ws_get_t(ctx, "nm_cnt", (size_t)N, &b.cnt))) return rc;
    if ((rc = ws_get_t(ctx, "nm_res", (size_t)N, &b.res))) return rc;
    if ((rc = ws_get_t(ctx, "nm_SR", (size_t)N + 1, &b.SR))) return rc;
    if ((rc = ws_get_t(ctx, "nm_SC", nz, &b.SC))) return rc;
    if ((rc = ws_get_t(ctx, "nm_sumBG", (size_t)N, &b.sumBG))) return rc;
    const char *yn[3] = {"nm_y0", "nm_y1", "nm_y2"}, *ygn[2] = {"nm_yg0", "nm_yg1"};
    for (int q = 0; q < 3; q++)
        if ((rc = ws_get_t(ctx, yn[q], (size_t)N, &b.y[q]))) return rc;
    for (int q = 0; q < 2; q++)
        if ((rc = ws_get_t(ctx, ygn[q], (size_t)N, &b.yg[q]))) return rc;
    if ((rc = ws_get_t(ctx, "nm_part", (size_t)4 * NM_BLOCKS_MAX, &b.part))) return rc;
    if ((rc = ws_get(ctx, "nm_state", nm_state_bytes(), &b.state))) return rc;
    if ((rc = ws_get_t(ctx, "nm_same", 1, &b.same))) return rc;
    if ((rc = ws_get_t(ctx, "nm_groups", (size_t)N, &b.groups))) return rc;
    if ((rc = ws_get_t(ctx, "nm_gout", (size_t)N, &b.gout))) return rc;
    if ((rc = ws_get_t(ctx, "nm_side", (size_t)N, &b.side))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = pin_grow(ctx, 2, 2 * EC_STATE_MAX_BYTES))) return rc;
    b.state_host = ctx->pin[2];
    b.ev[0] = ctx->ecev[0];
    b.ev[1] = ctx->ecev[1];
    std::vector<int32_t> rh;
    if (g->on_device) {
        rh.resize((size_t)N + 1);
        if ((rc = copy_d2h(ctx, rh.data(), g->R, sizeof(int32_t) * ((size_t)N + 1)))) return rc;
        b.R_host = rh.data();
    } else {
        b.R_host = g->R;
    }
    int mods = 0, rounds = 0;
    int64_t its = 0;
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(ctx, run_modularity(R, C, N, g->nnz, 10000, 1e-6, b, labels, Q, &mods, &its, &rounds, ctx->stream));
    ctx->tm.modularity_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->tm.modularity_iters = its;
    if (n_modules) *n_modules = mods;
    return BNA_OK;
}

int bna_cp(bna_ctx *ctx, const bna_csr *g, float *cp, double *mean)
{
    if (!ctx || !g || !cp) return fail(ctx, BNA_ERR_ARG, "bna_cp: NULL argument");
    if (g->N < 1 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_cp: empty graph");
    if (g->N > CP_MAX_N) return fail(ctx, BNA_ERR_ARG, "bna_cp: N above 524288 (the row kernel's LDS bitmap)");
    // the row kernel finds w_ik by rank in row i, so a host CSR must list
    // every row's columns strictly ascending (device CSRs built here do)
    if (!g->on_device && bna_check_csr(g, 1) != BNA_OK)
        return fail(ctx, BNA_ERR_FORMAT, "bna_cp: columns must ascend strictly within every row");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    float *d, *w = nullptr;
    unsigned long long *key = nullptr;
    if ((rc = ws_get_t(ctx, "cp", (size_t)g->N, &d))) return rc;
    if (V && g->nnz > 0) {
        if ((rc = ws_get_t(ctx, "cp_w", (size_t)g->nnz, &w))) return rc;
        if ((rc = ws_get_t(ctx, "cp_key", (size_t)1, &key))) return rc;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    HIPCHK(ctx, launch_cp(R, C, V, g->N, g->nnz, w, key, d, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    if ((rc = copy_d2h(ctx, cp, d, sizeof(float) * (size_t)g->N))) return rc;  // synchronises
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.cp_ms = ms;
    if (mean) {
        double s = 0.0;
        for (int i = 0; i < g->N; ++i) s += std::fabs((double)cp[i]);
        *mean = (double)((float)s / (float)g->N);
    }
    return BNA_OK;
}

int bna_lp(bna_ctx *ctx, const bna_csr *g, double *lp, int64_t *hist, int hist_len)
{
    if (!ctx || !g || !lp) return fail(ctx, BNA_ERR_ARG, "bna_lp: NULL argument");
    if (g->N < 2 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_lp: fewer than 2 vertices");
    if (hist_len < 0 || (hist_len > 0 && !hist)) return fail(ctx, BNA_ERR_ARG, "bna_lp: bad histogram buffer");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    const int N = g->N;
    void *vis, *fa, *fb;
    unsigned long long *cnt;
    if ((rc = ws_get(ctx, "lp_visited", (size_t)N * 32, &vis)) || (rc = ws_get(ctx, "lp_front", (size_t)N * 32, &fa)) ||
        (rc = ws_get(ctx, "lp_next", (size_t)N * 32, &fb)) || (rc = ws_get_t(ctx, "lp_count", 1, &cnt)))
        return rc;
    // ordered pairs at each distance, over every batch of 256 sources
    std::vector<int64_t> at(2, 0);
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    for (int s0 = 0; s0 < N; s0 += 256) {
        const int ns = std::min(256, N - s0);
        HIPCHK(ctx, launch_msbfs_init(N, s0, ns, vis, fa, ctx->stream));
        void *f = fa, *nx = fb;
        for (int d = 1; d < N; d++) {
            HIPCHK(ctx, launch_zero_u64(cnt, 1, ctx->stream));
            HIPCHK(ctx, launch_msbfs_level(R, C, N, f, nx, vis, ns, cnt, ctx->stream));
            unsigned long long c = 0;
            if ((rc = copy_d2h(ctx, &c, cnt, sizeof(c)))) return rc;
            if (c == 0) break;
            if ((int)at.size() <= d) at.resize((size_t)d + 1, 0);
            at[(size_t)d] += (int64_t)c;
            std::swap(f, nx);
        }
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[9]));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.lp_ms = ms;
    // CUBFS_LP.cpp:33-46: the sum of 1.0f / d over the pairs with d > 0, in
    // double, divided by N and by N - 1, and inverted
    double sum = 0.0;
    int64_t reached = 0;
    for (size_t d = 1; d < at.size(); d++) {
        sum += (double)at[d] * (double)(1.0f / (float)d);
        reached += at[d];
    }
    double r = sum;
    r /= N;
    r /= N - 1;
    *lp = 1.0 / r;
    if (hist_len > 0) {
        for (int d = 0; d < hist_len; d++) hist[d] = d < (int)at.size() ? at[(size_t)d] : 0;
        hist[0] = (int64_t)N * (N - 1) - reached;  // ordered pairs with no path
    }
    return BNA_OK;
}

int bna_lp_weighted(bna_ctx *ctx, const bna_csr *g, double *lp, float *eff, int *isolated)
{
    if (!ctx || !g || !lp) return fail(ctx, BNA_ERR_ARG, "bna_lp_weighted: NULL argument");
    if (g->N < 2 || !g->R || (g->nnz > 0 && !g->C))
        return fail(ctx, BNA_ERR_ARG, "bna_lp_weighted: fewer than 2 vertices");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    const int N = g->N;
    const int Np = round_up(N, FW_B);
    unsigned *bad;
    float *effd;
    double *rows;
    if ((rc = ws_get_t(ctx, "lpw_bad", 1, &bad)) || (rc = ws_get_t(ctx, "lpw_eff", (size_t)N, &effd)) ||
        (rc = ws_get_t(ctx, "lpw_rows", (size_t)N, &rows)))
        return rc;
    // the Np x Np cost matrix lives only for this call
    const size_t dbytes = (size_t)Np * Np * sizeof(float);
    float *D = nullptr;
    if ((rc = dev_alloc(ctx, (void **)&D, dbytes))) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    rc = BNA_OK;
    hipError_t e = launch_zero_u64(bad, 1, ctx->stream);
    if (e == hipSuccess) e = launch_fw_init(R, C, V, N, Np, D, bad, ctx->stream);
    unsigned hbad = 0;
    if (e != hipSuccess) rc = fail(ctx, BNA_ERR_HIP, std::string("fw init: ") + hipGetErrorString(e));
    if (!rc) rc = copy_d2h(ctx, &hbad, bad, sizeof(hbad));  // synchronises
    if (!rc && (hbad & 1u))
        rc = fail(ctx, BNA_ERR_ARG, "bna_lp_weighted: a weight is negative or NaN (costs are 1 / w)");
    if (!rc && (hbad & 2u)) rc = fail(ctx, BNA_ERR_FORMAT, "CSR: column out of range");
    for (int K = 0; !rc && K < Np / FW_B; K++)
        if ((e = launch_fw_block(D, Np, K, ctx->stream)) != hipSuccess)
            rc = fail(ctx, BNA_ERR_HIP, std::string("fw block: ") + hipGetErrorString(e));
    if (!rc && (e = launch_fw_rows(D, Np, N, effd, rows, ctx->stream)) != hipSuccess)
        rc = fail(ctx, BNA_ERR_HIP, std::string("fw rows: ") + hipGetErrorString(e));
    dev_free(ctx, D, dbytes);
    if (rc) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    std::vector<double> hr((size_t)N);
    std::vector<float> he((size_t)N);
    if ((rc = copy_d2h(ctx, hr.data(), rows, sizeof(double) * (size_t)N))) return rc;
    if ((rc = copy_d2h(ctx, he.data(), effd, sizeof(float) * (size_t)N))) return rc;
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.lpw_ms = ms;
    // CUBFW_Lp.cpp:111-113: the double sum over the ordered pairs, divided by N
    // and by N - 1, inverted; the rows' sums are added in row order
    double sum = 0.0;
    int iso = 0;
    for (int u = 0; u < N; u++) {
        sum += hr[(size_t)u];
        iso += he[(size_t)u] == 0.f;
    }
    sum /= N;
    sum /= N - 1;
    *lp = 1.0 / sum;
    if (eff) memcpy(eff, he.data(), sizeof(float) * (size_t)N);
    if (isolated) *isolated = iso;
    return BNA_OK;
}

int bna_bc(bna_ctx *ctx, const bna_csr *g, float *bc)
{
    if (!ctx || !g || !bc) return fail(ctx, BNA_ERR_ARG, "bna_bc: NULL argument");
    if (g->N < 1 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_bc: empty graph");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    const int N = g->N;
    // sources per batch: two workgroups per CU, at most ~2 GB of per-source state
    const int64_t per_src = (int64_t)N * 36 + 8;
    const int B = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)2 * ctx->ncu, (int64_t)N,
                                                              ((int64_t)2 << 30) / per_src}));
    int32_t *dist, *queue, *lvl;
    double *sigma, *delta, *out, *acc;
    if ((rc = ws_get_t(ctx, "bc_dist", (size_t)B * N, &dist)) || (rc = ws_get_t(ctx, "bc_queue", (size_t)B * N, &queue)) ||
        (rc = ws_get_t(ctx, "bc_lvl", (size_t)B * (N + 2), &lvl)) ||
        (rc = ws_get_t(ctx, "bc_sigma", (size_t)B * N, &sigma)) ||
        (rc = ws_get_t(ctx, "bc_delta", (size_t)B * N, &delta)) || (rc = ws_get_t(ctx, "bc_out", (size_t)B * N, &out)) ||
        (rc = ws_get_t(ctx, "bc_acc", (size_t)N, &acc)))
        return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    HIPCHK(ctx, launch_zero_u64(acc, N, ctx->stream));
    for (int s0 = 0; s0 < N; s0 += B) {
        const int ns = std::min(B, N - s0);
        HIPCHK(ctx, launch_bc_batch(R, C, N, s0, ns, dist, queue, lvl, sigma, delta, out, acc, ctx->stream));
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    std::vector<double> h((size_t)N);
    if ((rc = copy_d2h(ctx, h.data(), acc, sizeof(double) * (size_t)N))) return rc;  // synchronises
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.bc_ms = ms;
    for (int v = 0; v < N; v++) bc[v] = (float)h[(size_t)v];
    return BNA_OK;
}

int bna_eigcent(bna_ctx *ctx, const bna_csr *g, double tol, int max_iter, int output, double *ec, double *lambda,
                int *iters)
{
    if (!ctx || !g || !ec || !lambda) return fail(ctx, BNA_ERR_ARG, "bna_eigcent: NULL argument");
    if (g->N < 1 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_eigcent: empty graph");
    if (output != BNA_EC_WEIGHTED_REF && output != BNA_EC_BINARY_REF) return fail(ctx, BNA_ERR_ARG, "bad output mode");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int32_t *R, *C;
    const float *V;
    int rc = device_view(ctx, g, &R, &C, &V);
    if (rc) return rc;
    const int N = g->N;
    const EcPlan plan = ec_plan(N, g->nnz, ctx->ncu);
    EcBlocked bl;
    if (plan.blocked) {
        if ((rc = ws_get_t(ctx, "ec_qc", (size_t)N * plan.S, &bl.qc))) return rc;
        if ((rc = ws_get_t(ctx, "ec_qoff", (size_t)plan.items + 1, &bl.qoff))) return rc;
        if ((rc = ws_get_t(ctx, "ec_btot", (size_t)plan.G, &bl.btot))) return rc;
        if ((rc = ws_get_t(ctx, "ec_cq", (size_t)plan.quads * 4, &bl.cq))) return rc;
        if (V && (rc = ws_get_t(ctx, "ec_vq", (size_t)plan.quads * 4, &bl.vq))) return rc;
    }
    double *xa, *xb, *part, *outd;
    void *stt;
    if ((rc = ws_get_t(ctx, "ec_xa", (size_t)N, &xa))) return rc;
    if ((rc = ws_get_t(ctx, "ec_xb", (size_t)N, &xb))) return rc;
    if ((rc = ws_get_t(ctx, "ec_out", (size_t)N, &outd))) return rc;
    if ((rc = ws_get_t(ctx, "ec_part", (size_t)3 * ec_nparts(N), &part))) return rc;
    if ((rc = ws_get(ctx, "ec_state", ec_state_bytes(), &stt))) return rc;
    int it = 0;
    // the iteration's state crosses through pinned memory: make sure no earlier
    // copy through that buffer is still in flight before it is reused
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = pin_grow(ctx, 2, 2 * EC_STATE_MAX_BYTES))) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    HIPCHK(ctx, run_eigcent(R, C, V, N, plan, bl, tol, max_iter, output, xa, xb, part, stt, ctx->pin[2], outd,
                            lambda, &it, ctx->graphs, ctx->ecev, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    int rc2 = copy_d2h(ctx, ec, outd, sizeof(double) * (size_t)N);  // through pinned memory; synchronises
    if (rc2) return rc2;
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.ec_ms = ms;
    ctx->tm.ec_iters = it;
    if (iters) *iters = it;
    return BNA_OK;
}

int bna_eigcent_noalloc(bna_ctx *ctx, const bna_csr *g, double tol, int max_iter, int output, double *ec,
                        double *lambda, int *iters)
{
    if (!ctx || !g || !ec || !lambda) return fail(ctx, BNA_ERR_ARG, "bna_eigcent_noalloc: NULL argument");
    if (!g->on_device)
        return fail(ctx, BNA_ERR_ARG, "bna_eigcent_noalloc: g must be a device CSR (a host CSR needs an upload buffer)");
    if (g->N < 1 || !g->R || (g->nnz > 0 && !g->C)) return fail(ctx, BNA_ERR_ARG, "bna_eigcent_noalloc: empty graph");
    if (g->N > BNA_NOALLOC_EC_MAX_N)
        return fail(ctx, BNA_ERR_ARG, "bna_eigcent_noalloc: N = " + std::to_string(g->N) +
                                          " exceeds the scratch maximum of " + std::to_string(BNA_NOALLOC_EC_MAX_N) +
                                          " voxels (use bna_eigcent)");
    if (output != BNA_EC_WEIGHTED_REF && output != BNA_EC_BINARY_REF) return fail(ctx, BNA_ERR_ARG, "bad output mode");
    const int dev = ctx->device;
    if (dev >= NA_MAX_DEVICES) return fail(ctx, BNA_ERR_ARG, "bna_eigcent_noalloc: device ordinal too large");
    const int N = g->N;
    if ((size_t)3 * ec_nparts(N) * sizeof(double) > NA_PART_BYTES || ec_state_bytes() > EC_STATE_MAX_BYTES)
        return fail(ctx, BNA_ERR_ARG, "bna_eigcent_noalloc: scratch layout does not fit N");
    HIPCHK(ctx, hipSetDevice(dev));
    scratch_setup(dev);  // bna_create did the lookup already; this covers nothing else
    if (!na_ptr[dev])
        return fail(ctx, BNA_ERR_HIP, std::string("bna_eigcent_noalloc: hipGetSymbolAddress(scratch): ") +
                                          hipGetErrorString(na_err[dev]));
    unsigned char *base = (unsigned char *)na_ptr[dev];
    double *xa = (double *)base;
    double *xb = (double *)(base + NA_X_BYTES);
    double *part = (double *)(base + 2 * NA_X_BYTES);
    void *stt = base + 2 * NA_X_BYTES + NA_PART_BYTES;
    const EcPlan plan;  // the plain SpMV: the blocked layout needs a copy of the graph
    // the host side of the stopping state (two slots): pageable, read after
    // the copy into it has completed
    alignas(16) unsigned char hstate[2 * EC_STATE_MAX_BYTES];
    std::lock_guard<std::mutex> lk(na_call_mu[dev]);
    int it = 0;
    HIPCHK(ctx, hipEventRecord(ctx->ev[8], ctx->stream));
    HIPCHK(ctx, run_eigcent(g->R, g->C, g->V, N, plan, EcBlocked{}, tol, max_iter, output, xa, xb, part, stt, hstate,
                            ec, lambda, &it, false, ctx->ecev, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[9], ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // ec written, scratch free for the next call
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[8], ctx->ev[9]));
    ctx->tm.ec_ms = ms;
    ctx->tm.ec_iters = it;
    if (iters) *iters = it;
    return BNA_OK;
}

}  // extern "C"

// ------------------------------------------------------------ autotuner
// The tile kernel claims tiles in table order, each XCD from its own contiguous
// range, so the band height G of the table (tiles of G consecutive tile-rows
// for each tile-column in turn) fixes which operand panels the workgroups of an
// XCD share in L2.  The best G depends on the series length (the panel bytes per
// tile) and the precision (bytes per point).  bna_autotune times the emission
// pass for every candidate on a synthetic subject of the bucket's length and
// keeps the fastest; results persist in a small text file, one line per
// (precision, bucket): "<precision> <bucket> <G> <ms>".
extern "C" {

int bna_autotune_get(bna_ctx *ctx, int precision, int L, int *band_rows)
{
    if (!ctx || !band_rows || precision < BNA_FP32 || precision > BNA_BF16X3 || L < 2)
        return fail(ctx, BNA_ERR_ARG, "bna_autotune_get: bad argument");
    *band_rows = ctx->band[precision][l_bucket(std::max(LPAD_MIN, round_up(L, KPAD)))];
    return BNA_OK;
}

int bna_autotune_set(bna_ctx *ctx, int precision, int L, int band_rows)
{
    if (!ctx || precision < BNA_FP32 || precision > BNA_BF16X3 || L < 2 || band_rows < 1 || band_rows > 1024)
        return fail(ctx, BNA_ERR_ARG, "bna_autotune_set: bad argument");
    ctx->band[precision][l_bucket(std::max(LPAD_MIN, round_up(L, KPAD)))] = band_rows;
    ctx->tt_nb = -1;
    return BNA_OK;
}

int bna_autotune_load(bna_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, BNA_ERR_ARG, "bna_autotune_load: NULL argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(ctx, BNA_ERR_IO, std::string("bna_autotune_load: cannot open ") + path);
    int p, b, g, n = 0;
    double ms;
    while (std::fscanf(f, "%d %d %d %lf", &p, &b, &g, &ms) == 4) {
        if (p < BNA_FP32 || p > BNA_BF16X3 || b < 0 || b > 3 || g < 1 || g > 1024) {
            std::fclose(f);
            return fail(ctx, BNA_ERR_IO, std::string("bna_autotune_load: bad line in ") + path);
        }
        ctx->band[p][b] = g;
        n++;
    }
    std::fclose(f);
    ctx->tt_nb = -1;
    return n > 0 ? BNA_OK : fail(ctx, BNA_ERR_IO, std::string("bna_autotune_load: no entries in ") + path);
}

int bna_autotune(bna_ctx *ctx, int precision, int L, int N, const char *cache_path, int *band_rows, double *best_ms)
{
    if (!ctx || precision < BNA_FP32 || precision > BNA_BF16X3 || L < 2 || N < TILE)
        return fail(ctx, BNA_ERR_ARG, "bna_autotune: bad argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int Lp = std::max(LPAD_MIN, round_up(L, KPAD));
    const int b = l_bucket(Lp);
    std::vector<float> x((size_t)N * L);
    int rc = bna_synth_factor_model(x.data(), N, L, 20, 0.6, 0xB17A7000ull + (uint64_t)b, BNA_TIME_MAJOR);
    if (rc) return fail(ctx, rc, "bna_autotune: synthetic series");
    bna_series *s = nullptr;
    if ((rc = bna_series_create(ctx, x.data(), N, L, BNA_TIME_MAJOR, 0, &s))) return rc;
    bna_opts o;
    bna_opts_default(&o);
    o.thr_type = BNA_THR_R;
    o.thr = 0.4f;
    o.precision = precision;
    o.weighted = 0;
    const int saved = ctx->band[precision][b];
    int best = saved;
    double best_t = 1e30;
    for (int c = 0; c < 4 && rc == BNA_OK; c++) {
        ctx->band[precision][b] = BAND_CANDIDATES[c];
        double t = 1e30;
        for (int rep = 0; rep < 3 && rc == BNA_OK; rep++) {  // first run warms the table and workspaces
            bna_csr g;
            rc = bna_build_csr_series(ctx, s, &o, &g);
            if (rc == BNA_OK) {
                if (rep > 0) t = std::min(t, ctx->tm.corr_ms);
                bna_csr_release(ctx, &g);
            }
        }
        if (rc == BNA_OK && t < best_t) {
            best_t = t;
            best = BAND_CANDIDATES[c];
        }
    }
    bna_series_destroy(ctx, s);
    ctx->band[precision][b] = rc == BNA_OK ? best : saved;
    ctx->tt_nb = -1;
    if (rc) return rc;
    if (band_rows) *band_rows = best;
    if (best_ms) *best_ms = best_t;
    if (cache_path) {
        // keep the file's other entries, replace this one
        std::vector<std::string> keep;
        if (FILE *f = std::fopen(cache_path, "r")) {
            int p, bb, g;
            double ms;
            while (std::fscanf(f, "%d %d %d %lf", &p, &bb, &g, &ms) == 4)
                if (!(p == precision && bb == b)) {
                    char line[96];
                    std::snprintf(line, sizeof line, "%d %d %d %.6f", p, bb, g, ms);
                    keep.push_back(line);
                }
            std::fclose(f);
        }
        FILE *f = std::fopen(cache_path, "w");
        if (!f) return fail(ctx, BNA_ERR_IO, std::string("bna_autotune: cannot write ") + cache_path);
        for (const std::string &l : keep) std::fprintf(f, "%s\n", l.c_str());
        std::fprintf(f, "%d %d %d %.6f\n", precision, b, best, best_t);
        std::fclose(f);
    }
    return BNA_OK;
}

}  // extern "C"

// ------------------------------------------------------------ group averaging
// The reference's avg = y/b path (CorMat.cpp:418-670): per-pair sums of r or
// Fisher z over the subjects of a directory, averaged, then thresholded like a
// single subject's matrix.  The sums live in device memory for the whole group.
struct GroupCopy {
    float *xs = nullptr;  // voxel-major series, row stride ld
    double *mu = nullptr, *sd = nullptr;
    int64_t ld = 0;
    int L = 0;
    size_t xs_bytes = 0;
};
struct bna_group {
    int N = 0, fisher = 0, count = 0;
    bool finished = false;
    int64_t M1 = 0;
    float *acc = nullptr;  // packed upper triangle, row-major, i < j
    // exact averages: each subject's normalised series is kept so that the
    // pairs at the threshold (and every kept edge's value) can be recomputed
    // from FP64 r in the reference's float arithmetic
    bool exact = true;
    std::vector<GroupCopy> subs;
    double band_max = 0.0;       // largest MFMA error bound of the subjects' r
    uint32_t *flags = nullptr;   // Fisher sums: pairs with |r| > GROUP_FLAG_R in some subject
    int64_t last_fixed = 0;      // pairs the last bna_group_build_csr recomputed at the threshold
};
// Fisher sums: a pair whose |r| exceeds this in some subject is always
// recomputed (the slope of z = atanh r is 1 / (1 - r^2): 5.3 here).
constexpr float GROUP_FLAG_R = 0.9f;
static size_t group_flag_bytes(int64_t M1) { return (size_t)((M1 + 63) / 64) * 8; }  // whole 64-bit words

static inline int64_t packed_pos(int64_t N, int64_t i, int64_t j) { return i * N - (i + 1) * i / 2 + j - i - 1; }

static float key_to_float(uint32_t k)
{
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}

// k-th largest of the averages by a 4-pass radix select over order-preserving
// keys.  above: values strictly greater; ties: values equal to it.
static int group_select(bna_ctx *ctx, bna_group *g, int64_t k, float *r_star, int64_t *above, int64_t *ties)
{
    unsigned long long *hist = nullptr;
    int rc = ws_get_t(ctx, "group_hist", 256, &hist);
    if (rc) return rc;
    uint32_t prefix = 0, mask = 0;
    int64_t need = k, gt = 0, last = 0;
    std::vector<unsigned long long> h(256);
    for (int shift = 24; shift >= 0; shift -= 8) {
        HIPCHK(ctx, hipMemsetAsync(hist, 0, 256 * sizeof(unsigned long long), ctx->stream));
        HIPCHK(ctx, launch_group_digits(g->acc, g->M1, prefix, mask, shift, hist, ctx->stream));
        if ((rc = copy_d2h(ctx, h.data(), hist, 256 * sizeof(unsigned long long)))) return rc;
        int d = 255;
        for (; d >= 0; d--) {
            if ((int64_t)h[d] >= need) break;
            need -= (int64_t)h[d];
            gt += (int64_t)h[d];
        }
        if (d < 0) return fail(ctx, BNA_ERR_DENSITY, "group density: fewer pairs than the edge count");
        prefix |= (uint32_t)d << shift;
        mask |= 0xffu << shift;
        last = (int64_t)h[d];
    }
    *r_star = key_to_float(prefix);
    *above = gt;
    *ties = last;
    return BNA_OK;
}

// Bound on |device average - exact average| for a pair the exact path does not
// recompute: each subject's term differs by the MFMA error band (through the
// slope of atanh for Fisher sums, whose pairs with |r| > GROUP_FLAG_R in some
// subject are always recomputed) plus the float rounding of r; both sides round
// their S float additions (partial sums at most S, or S z_max), the division by
// S and, for Fisher sums, the inverse transform.  u = 2^-24 is half an ulp of 1.
static double group_bound(const bna_group *g)
{
    const double u = std::ldexp(1.0, -24);
    const double S = (double)g->count, b = g->band_max;
    if (!g->fisher) return 1.5 * (b + u / 2 + 2.0 * S * u + 2.0 * u);
    const double R = (double)GROUP_FLAG_R + b + u;  // |r64| of an unflagged pair
    const double slope = 1.0 / (1.0 - R * R);
    const double zmax = std::atanh(R) + 1e-6;
    const double term = slope * (b + u / 2) + 16.0 * u * (1.0 + zmax);  // logf and the quotient, both sides
    return 1.5 * (term + 2.0 * S * zmax * u + 2.0 * zmax * u + 16.0 * u);
}

// The pairs listed in pairs[0, *count) get their exact averages (the count is
// read on the device, so the list may come straight from a kernel).
static int group_exact_pairs(bna_ctx *ctx, bna_group *g, const int2 *pairs, const unsigned long long *count,
                             int64_t cap)
{
    std::vector<GroupSubj> hs(g->subs.size());
    for (size_t q = 0; q < g->subs.size(); q++) {
        const GroupCopy &c = g->subs[q];
        hs[q] = GroupSubj{c.xs, c.ld, c.mu, c.sd, c.L, 0};
    }
    GroupSubj *ds = nullptr;
    int rc = ws_get_t(ctx, "group_subjs", hs.size(), &ds);
    if (rc) return rc;
    if ((rc = h2d_pinned(ctx, 0, ds, hs.data(), hs.size() * sizeof(GroupSubj)))) return rc;
    HIPCHK(ctx, launch_group_exact(pairs, count, cap, ds, (int)hs.size(), g->fisher, g->acc, g->N, ctx->stream));
    return BNA_OK;
}

// Recomputes the averages of every pair within w of t (and every flagged pair).
static int group_fix_band(bna_ctx *ctx, bna_group *g, float t, float w, int64_t *n_fixed)
{
    unsigned long long *cnt = nullptr;
    int rc = ws_get_t(ctx, "group_pair_count", 1, &cnt);
    if (rc) return rc;
    int64_t cap = std::max<int64_t>(1 << 20, (int64_t)(ctx->ws["group_pairs"].bytes / sizeof(int2)));
    for (int attempt = 0; attempt < 2; attempt++) {
        int2 *pairs = nullptr;
        if ((rc = ws_get_t(ctx, "group_pairs", (size_t)cap, &pairs))) return rc;
        HIPCHK(ctx, launch_zero_u64(cnt, 1, ctx->stream));
        HIPCHK(ctx, launch_group_cands(g->acc, g->N, t, w, g->flags, pairs, cnt, cap, ctx->stream));
        unsigned long long n = 0;
        if ((rc = copy_d2h(ctx, &n, cnt, sizeof(n)))) return rc;
        if ((int64_t)n <= cap) {
            *n_fixed = (int64_t)n;
            return n ? group_exact_pairs(ctx, g, pairs, cnt, cap) : BNA_OK;
        }
        cap = (int64_t)n + (int64_t)n / 8 + 1024;
    }
    return fail(ctx, BNA_ERR_DEVICE_NOMEM, "group: band pair list keeps overflowing");
}

extern "C" {

int bna_group_create(bna_ctx *ctx, int N, int fisher, bna_group **out)
{
    if (!ctx || !out) return fail(ctx, BNA_ERR_ARG, "bna_group_create: NULL argument");
    *out = nullptr;
    if (N < 2) return fail(ctx, BNA_ERR_ARG, "bna_group_create: N < 2");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    bna_group *g = new (std::nothrow) bna_group();
    if (!g) return fail(ctx, BNA_ERR_NOMEM, "bna_group_create: host allocation");
    g->N = N;
    g->fisher = fisher ? 1 : 0;
    g->M1 = (int64_t)N * (N - 1) / 2;
    int rc = dev_alloc(ctx, (void **)&g->acc, (size_t)g->M1 * 4);
    if (rc) {
        delete g;
        return rc;
    }
    hipError_t e = hipMemsetAsync(g->acc, 0, (size_t)g->M1 * 4, ctx->stream);
    if (e != hipSuccess) {
        dev_free(ctx, g->acc, (size_t)g->M1 * 4);
        delete g;
        return fail(ctx, BNA_ERR_HIP, std::string("bna_group_create: ") + hipGetErrorString(e));
    }
    *out = g;
    return BNA_OK;
}

int bna_group_destroy(bna_ctx *ctx, bna_group *g)
{
    if (!g) return BNA_OK;
    if (!ctx) return fail(nullptr, BNA_ERR_ARG, "bna_group_destroy: ctx is NULL");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dev_free(ctx, g->acc, (size_t)g->M1 * 4);
    for (GroupCopy &c : g->subs) {
        dev_free(ctx, c.xs, c.xs_bytes);
        dev_free(ctx, c.mu, (size_t)g->N * 8);
        dev_free(ctx, c.sd, (size_t)g->N * 8);
    }
    if (g->flags) dev_free(ctx, g->flags, group_flag_bytes(g->M1));
    (void)hipStreamSynchronize(ctx->stream);
    delete g;
    return BNA_OK;
}

int bna_group_set_exact(bna_ctx *ctx, bna_group *g, int exact)
{
    if (!ctx || !g) return fail(ctx, BNA_ERR_ARG, "bna_group_set_exact: NULL argument");
    if (g->count > 0) return fail(ctx, BNA_ERR_ARG, "bna_group_set_exact: subjects were already added");
    g->exact = exact != 0;
    return BNA_OK;
}

int bna_group_add(bna_ctx *ctx, bna_group *g, bna_series *s, int precision)
{
    if (!ctx || !g || !s) return fail(ctx, BNA_ERR_ARG, "bna_group_add: NULL argument");
    if (g->finished) return fail(ctx, BNA_ERR_ARG, "bna_group_add: the group is already averaged");
    if (s->N != g->N)
        return fail(ctx, BNA_ERR_ARG, "bna_group_add: subject has " + std::to_string(s->N) + " voxels, group " +
                                          std::to_string(g->N));
    if (precision < BNA_FP32 || precision > BNA_BF16X3) return fail(ctx, BNA_ERR_ARG, "bna_group_add: bad precision");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    memset(&ctx->tm, 0, sizeof(ctx->tm));
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    int rc;
    if (s->prec_ready != precision && (rc = series_normalize(ctx, s, precision))) return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    SubjDesc h{};
    h.a0 = s->panel;
    h.N = s->N;
    h.N_pad = s->N_pad;
    h.L_pad = s->L_pad;
    h.nb = s->N_pad / TILE;
    h.dense = g->acc;
    h.fisher = g->fisher;
    if (g->exact && g->fisher) {
        if (!g->flags) {
            if ((rc = dev_alloc(ctx, (void **)&g->flags, group_flag_bytes(g->M1)))) return rc;
            HIPCHK(ctx, launch_zero_u64(g->flags, (int64_t)(group_flag_bytes(g->M1) / 8), ctx->stream));
        }
        h.flags = g->flags;
        h.flag_r = GROUP_FLAG_R;
    }
    SubjDesc *dsd = nullptr;
    if ((rc = ws_get_t(ctx, "group_subjdesc", 1, &dsd))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(dsd, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    TileItem *tiles = nullptr;
    int64_t nt = 0;
    const int64_t kstride = (int64_t)s->N_pad * panel_unit(precision);
    if ((rc = get_tile_table(ctx, h.nb, s->panel, kstride, s->L_pad / 16, ctx->band[precision][l_bucket(s->L_pad)],
                             &tiles, &nt)))
        return rc;
    unsigned int *claim = nullptr;
    if ((rc = ws_get_t(ctx, "tile_claim", 8, &claim))) return rc;
    const int grid = ctx->grid_limit > 0 ? std::min(ctx->grid, ctx->grid_limit) : ctx->grid;
    HIPCHK(ctx, launch_corr(precision, 2, dsd, tiles, nt, claim, nullptr, grid, ctx->stream));
    if (g->exact) {
        // the subject's series as the FP64 recheck reads it, with its mean and norm
        GroupCopy c;
        c.ld = s->xv ? s->L_pad : s->L;
        c.L = s->L;
        c.xs_bytes = (size_t)s->N * (size_t)c.ld * 4;
        if ((rc = dev_alloc(ctx, (void **)&c.xs, c.xs_bytes)) || (rc = dev_alloc(ctx, (void **)&c.mu, (size_t)s->N * 8)) ||
            (rc = dev_alloc(ctx, (void **)&c.sd, (size_t)s->N * 8))) {
            dev_free(ctx, c.xs, c.xs_bytes);
            dev_free(ctx, c.mu, (size_t)s->N * 8);
            return rc;
        }
        HIPCHK(ctx, hipMemcpyAsync(c.xs, s->xv ? s->xv : s->x, c.xs_bytes, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(c.mu, s->mu, (size_t)s->N * 8, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(c.sd, s->sd, (size_t)s->N * 8, hipMemcpyDeviceToDevice, ctx->stream));
        g->subs.push_back(c);
        g->band_max = std::max(g->band_max, band_for(precision, s->L_pad));
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[2]));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->tm.znorm_ms = ms;
    if (hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]) == hipSuccess) ctx->tm.corr_ms = ms;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[2]) == hipSuccess) ctx->tm.total_ms = ms;
    ctx->tm.corr_launches = 1;
    g->count++;
    return BNA_OK;
}

int bna_group_finish(bna_ctx *ctx, bna_group *g)
{
    if (!ctx || !g) return fail(ctx, BNA_ERR_ARG, "bna_group_finish: NULL argument");
    if (g->finished) return BNA_OK;
    if (g->count < 1) return fail(ctx, BNA_ERR_ARG, "bna_group_finish: no subject was added");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_group_finish(g->acc, g->M1, g->count, g->fisher, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    g->finished = true;
    return BNA_OK;
}

int bna_group_count(const bna_group *g) { return g ? g->count : 0; }

int64_t bna_group_band_pairs(const bna_group *g) { return g ? g->last_fixed : -1; }

int bna_group_download(bna_ctx *ctx, const bna_group *g, float *packed)
{
    if (!ctx || !g || !packed) return fail(ctx, BNA_ERR_ARG, "bna_group_download: NULL argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return copy_d2h(ctx, packed, g->acc, (size_t)g->M1 * 4);
}

int bna_build_csr_into(bna_ctx *ctx, bna_series *s, const bna_opts *o, int32_t *R, int32_t *C, float *V, int64_t cap,
                       int64_t *nnz)
{
    if (!ctx || !s || !o || !nnz) return fail(ctx, BNA_ERR_ARG, "bna_build_csr_into: NULL argument");
    *nnz = 0;
    bna_csr g{};
    int rc = bna_build_csr_series(ctx, s, o, &g);
    if (rc) return rc;
    *nnz = g.nnz;
    if (!R) return bna_csr_release(ctx, &g);  // size query
    if (!C || cap < g.nnz || (g.V && !V)) {
        (void)bna_csr_release(ctx, &g);
        return fail(ctx, BNA_ERR_OVERFLOW, "bna_build_csr_into: " + std::to_string(g.nnz) +
                                               " entries, capacity " + std::to_string(cap));
    }
    rc = copy_d2h(ctx, R, g.R, sizeof(int32_t) * ((size_t)g.N + 1));
    if (!rc && g.nnz) rc = copy_d2h(ctx, C, g.C, sizeof(int32_t) * (size_t)g.nnz);
    if (!rc && g.nnz && g.V) rc = copy_d2h(ctx, V, g.V, sizeof(float) * (size_t)g.nnz);
    const int rr = bna_csr_release(ctx, &g);
    return rc ? rc : rr;
}

int bna_corr_packed(bna_ctx *ctx, bna_series *s, int precision, float *packed)
{
    if (!ctx || !s || !packed) return fail(ctx, BNA_ERR_ARG, "bna_corr_packed: NULL argument");
    bna_group *g = nullptr;
    int rc = bna_group_create(ctx, s->N, 0, &g);
    if (rc) return rc;
    g->exact = false;  // one subject's r as the tile kernel gives it
    if (!(rc = bna_group_add(ctx, g, s, precision)) && !(rc = bna_group_finish(ctx, g)))
        rc = bna_group_download(ctx, g, packed);
    bna_group_destroy(ctx, g);
    return rc;
}

int bna_select_density(bna_ctx *ctx, bna_series *s, int precision, double density_pct, double *r_star, int64_t *k)
{
    if (!ctx || !s || !r_star) return fail(ctx, BNA_ERR_ARG, "bna_select_density: NULL argument");
    bna_opts o;
    bna_opts_default(&o);
    o.thr_type = BNA_THR_DENSITY;
    o.thr = density_pct;
    o.weighted = 0;
    o.precision = precision;
    bna_csr g{};
    int rc = bna_build_csr_series(ctx, s, &o, &g);
    if (rc) return rc;
    *r_star = g.thr_used;
    if (k) *k = g.nnz / 2;
    return bna_csr_release(ctx, &g);
}

int bna_group_build_csr(bna_ctx *ctx, bna_group *g, const bna_opts *o, bna_csr *out)
{
    if (!ctx || !g || !o || !out) return fail(ctx, BNA_ERR_ARG, "bna_group_build_csr: NULL argument");
    memset(out, 0, sizeof(*out));
    if (!g->finished) return fail(ctx, BNA_ERR_ARG, "bna_group_build_csr: call bna_group_finish first");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int N = g->N;
    const bool dens = o->thr_type == BNA_THR_DENSITY;
    if (!dens && o->thr_type != BNA_THR_R) return fail(ctx, BNA_ERR_ARG, "bna_group_build_csr: bad threshold type");
    float thr = (float)o->thr;
    int64_t p_cut = 0;
    int rc;
    memset(&ctx->tm, 0, sizeof(ctx->tm));
    const auto t0 = std::chrono::steady_clock::now();
    const bool exact = g->exact && g->count > 0 && (int)g->subs.size() == g->count;
    g->last_fixed = 0;
    if (exact) {
        // Pairs whose device average may sit on the other side of the threshold
        // than the exact one get the exact value.  For a density, the k-th
        // largest device average t lies within the bound B of the exact r*, so
        // recomputing every pair within 2B of t leaves only pairs whose order
        // against r* is certain; the selection below then runs on those values.
        const double B = group_bound(g);
        float t = thr;
        if (dens) {
            const int64_t k = bna_density_to_k(N, (float)o->thr);
            if (k < 1) return fail(ctx, BNA_ERR_DENSITY, "density: the percentage gives no edge (k = 0)");
            if (k > g->M1) return fail(ctx, BNA_ERR_DENSITY, "density: more edges than pairs");
            int64_t a0 = 0, t0_ = 0;
            if ((rc = group_select(ctx, g, k, &t, &a0, &t0_))) return rc;
        }
        if ((rc = group_fix_band(ctx, g, t, (float)((dens ? 2.0 : 1.0) * B), &g->last_fixed))) return rc;
    }
    if (dens) {
        const int64_t k = bna_density_to_k(N, (float)o->thr);
        if (k < 1) return fail(ctx, BNA_ERR_DENSITY, "density: the percentage gives no edge (k = 0)");
        if (k > g->M1) return fail(ctx, BNA_ERR_DENSITY, "density: more edges than pairs");
        int64_t above = 0, ties = 0;
        if ((rc = group_select(ctx, g, k, &thr, &above, &ties))) return rc;
        const int64_t m = k - above;  // ties kept, the first m in row-major order (CorMat.cpp:784-812)
        p_cut = g->M1;
        if (m < ties) {
            int64_t *tr = nullptr;
            if ((rc = ws_get_t(ctx, "group_ties", (size_t)N, &tr))) return rc;
            HIPCHK(ctx, launch_group_ties(g->acc, N, thr, tr, ctx->stream));
            std::vector<int64_t> rows(N);
            if ((rc = copy_d2h(ctx, rows.data(), tr, sizeof(int64_t) * N))) return rc;
            int64_t before = 0;
            int i = 0;
            for (; i < N && before + rows[i] < m; i++) before += rows[i];
            if (i >= N - 1) return fail(ctx, BNA_ERR_DENSITY, "group density: tie counts are inconsistent");
            std::vector<float> row((size_t)(N - 1 - i));
            if ((rc = copy_d2h(ctx, row.data(), g->acc + packed_pos(N, i, i + 1), sizeof(float) * row.size())))
                return rc;
            int64_t want = m - before, c = 0;
            size_t q = 0;
            for (; q < row.size(); q++)
                if (row[q] == thr && ++c == want) break;
            if (q == row.size()) return fail(ctx, BNA_ERR_DENSITY, "group density: tie row is inconsistent");
            p_cut = packed_pos(N, i, i + 1 + (int64_t)q);
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    uint8_t *keep = nullptr;
    unsigned int *flag = nullptr;
    int64_t *cnt = nullptr;
    const int64_t ld = ((int64_t)N + 63) / 64 * 64;
    if ((rc = ws_get_t(ctx, "group_keep", (size_t)N * (size_t)ld, &keep))) return rc;
    if ((rc = ws_get_t(ctx, "group_flag", 1, &flag))) return rc;
    if ((rc = ws_get_t(ctx, "group_cnt", (size_t)N, &cnt))) return rc;
    HIPCHK(ctx, hipMemsetAsync(flag, 0, sizeof(unsigned int), ctx->stream));
    HIPCHK(ctx, launch_group_keep(g->acc, N, ld, thr, dens ? 1 : 0, p_cut, keep, flag, ctx->stream));
    HIPCHK(ctx, launch_group_rowcount(keep, N, ld, cnt, ctx->stream));
    std::vector<int64_t> c(N);
    unsigned int bad = 0;
    if ((rc = copy_d2h(ctx, c.data(), cnt, sizeof(int64_t) * N))) return rc;
    if ((rc = copy_d2h(ctx, &bad, flag, sizeof(bad)))) return rc;
    if (bad) return fail(ctx, BNA_ERR_FORMAT, "FormFullGraph: illegal correlation coefficient (an average above 1)");
    std::vector<int32_t> R((size_t)N + 1);
    int64_t nnz = 0;
    R[0] = 0;
    for (int i = 0; i < N; i++) {
        nnz += c[i];
        if (nnz > INT32_MAX) return fail(ctx, BNA_ERR_OVERFLOW, "group graph: more than 2^31 - 1 entries");
        R[i + 1] = (int32_t)nnz;
    }
    if (dens && nnz != 2 * bna_density_to_k(N, (float)o->thr))
        return fail(ctx, BNA_ERR_DENSITY, "group density: selected edge count differs from k");
    bna_csr r{};
    r.N = N;
    r.nnz = nnz;
    r.on_device = 1;
    if ((rc = out_alloc(ctx, (void **)&r.R, sizeof(int32_t) * ((size_t)N + 1)))) return rc;
    if ((rc = out_alloc(ctx, (void **)&r.C, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)))) {
        bna_csr_release(ctx, &r);
        return rc;
    }
    if (o->weighted && (rc = out_alloc(ctx, (void **)&r.V, sizeof(float) * (size_t)std::max<int64_t>(nnz, 1)))) {
        bna_csr_release(ctx, &r);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(r.R, R.data(), sizeof(int32_t) * R.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_group_fill(keep, ld, g->acc, N, r.R, r.C, r.V, ctx->stream);
    if (e == hipSuccess && exact && r.V && nnz > 0) {
        // every kept edge's value is the exact average
        int2 *pairs = nullptr;
        unsigned long long *cnt = nullptr;
        const int64_t cap = nnz / 2;
        if ((rc = ws_get_t(ctx, "group_pairs", (size_t)cap, &pairs)) ||
            (rc = ws_get_t(ctx, "group_pair_count", 1, &cnt))) {
            bna_csr_release(ctx, &r);
            return rc;
        }
        e = launch_zero_u64(cnt, 1, ctx->stream);
        if (e == hipSuccess) e = launch_csr_upper_pairs(r.R, r.C, N, pairs, cnt, cap, ctx->stream);
        if (e == hipSuccess && (rc = group_exact_pairs(ctx, g, pairs, cnt, cap))) {
            bna_csr_release(ctx, &r);
            return rc;
        }
        if (e == hipSuccess) e = launch_group_refill(r.R, r.C, g->acc, N, r.V, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        bna_csr_release(ctx, &r);
        return fail(ctx, BNA_ERR_HIP, std::string("bna_group_build_csr: ") + hipGetErrorString(e));
    }
    r.thr_used = thr;
    r.density_pct = 100.0 * (double)nnz / (double)g->M1 / 2.0;
    *out = r;
    const auto t2 = std::chrono::steady_clock::now();
    ctx->tm.select_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ctx->tm.csr_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    ctx->tm.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    return BNA_OK;
}

}  // extern "C"
