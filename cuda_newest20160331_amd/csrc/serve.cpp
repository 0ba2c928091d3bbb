// serve.cpp -- the request loop: subjects of any N and L are submitted from host
// threads, admitted by device-memory footprint, and built in rounds by worker
// threads that each own a context (a stream and a workspace).  A round is one
// bna_build_csr_batch over the requests admitted together, followed by each
// request's degree / eigenvector centrality and the copy of its results to
// the host.  With two or more workers a round can start while another is
// still on the device, so requests that arrive while the GPU is busy are
// admitted as soon as memory frees up and their kernels overlap.
//
// The reference processes subjects one at a time in a single host thread
// (CorMat.cpp:195-384); there is no counterpart of this loop there.
#include "bna.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Req {
    int64_t id = 0;
    std::vector<float> x;
    int N = 0, L = 0, layout = 0;
    bna_opts opts{};
    int want_degree = 0, want_ec = 0, device_result = 0;
    int64_t footprint = 0;
    Clock::time_point t_submit, t_start;
    bool done = false;
    bna_result res{};
};

}  // namespace

struct bna_server {
    bna_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> queue;
    std::map<int64_t, std::shared_ptr<Req>> all;  // queued, running or finished and not yet collected
    int64_t next_id = 1;
    int64_t reserved = 0;  // footprint of admitted, unfinished requests
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<bna_ctx *> ctxs;
    bna_server_stats st{};
    int busy_workers = 0;
    // Small rounds run ungated; while any is on the device, big rounds (issued
    // in pieces) hold back their next piece, so a small subject never queues
    // behind a whole long kernel of a big one.
    std::mutex gate_mu;
    std::condition_variable gate_cv;
    int small_active = 0;
    int grid_cap = 0;  // workgroup cap of big rounds' tile launches: the CUs minus one XCD's share
};

static void gate_fn(void *arg)
{
    bna_server *srv = (bna_server *)arg;
    std::unique_lock<std::mutex> lk(srv->gate_mu);
    srv->gate_cv.wait(lk, [&] { return srv->small_active == 0; });
}

// Tile-kernel time of a request, ms: tiles x L_pad at ~3e6 tile-points per ms,
// two passes for a density threshold.
static double est_ms(const Req &r)
{
    const double nb = (double)((r.N + 255) / 256), Lp = (double)std::max(64, (r.L + 15) / 16 * 16);
    return nb * (nb + 1) / 2.0 * Lp / 3.0e6 * (r.opts.thr_type == BNA_THR_DENSITY ? 2.0 : 1.0);
}

extern "C" int64_t bna_request_footprint(int N, int L, const bna_opts *o)
{
    if (N < 2 || L < 2 || !o) return -1;
    const int64_t Np = (N + 255) / 256 * 256, Lp = std::max<int64_t>(64, (L + 15) / 16 * 16), nb1 = Np / 256 + 1;
    const int unit = o->precision == BNA_BF16 ? 32 : 64;
    const double M1 = (double)N * (N - 1) / 2.0;
    // edges: the density target, or 2 % of the pairs for an r threshold
    double edges = o->thr_type == BNA_THR_DENSITY ? M1 * o->thr / 100.0 : 0.02 * M1;
    edges = std::max(edges, 65536.0);
    int64_t b = 0;
    b += 4 * (int64_t)N * L + 4 * (int64_t)N * Lp + 17 * (int64_t)N;  // series, voxel-major copy, mean/norm/valid
    b += (Lp / 16) * Np * unit;                                           // operand panel
    b += 32 * (int64_t)N * nb1;                                           // fragment tables, both orders
    b += (int64_t)(16.0 * edges * 1.15) + (1 << 23);                      // entries, candidates, pools
    b += 4 * (int64_t)(N + 1) + (int64_t)(8.0 * 2.0 * edges);              // the CSR
    return b;
}


// Requests of one precision, oldest first, that fit in what the budget has
// left; a request larger than the whole budget runs alone on an idle server.
static std::vector<std::shared_ptr<Req>> admit(bna_server *srv)
{
    std::vector<std::shared_ptr<Req>> take;
    const int64_t budget = (int64_t)srv->cfg.hbm_budget;
    int prec = -1;
    for (auto it = srv->queue.begin(); it != srv->queue.end() && (int)take.size() < srv->cfg.max_batch;) {
        const auto &r = *it;
        if (prec >= 0 && r->opts.precision != prec) {
            ++it;
            continue;
        }
        const bool alone = srv->reserved == 0 && take.empty();
        if (srv->reserved + r->footprint <= budget || alone) {
            prec = r->opts.precision;
            srv->reserved += r->footprint;
            take.push_back(r);
            it = srv->queue.erase(it);
            if (srv->reserved > budget) break;  // the oversized one runs by itself
        } else {
            break;  // keep arrival order: nothing overtakes a request waiting for memory
        }
    }
    return take;
}

static void run_round(bna_server *srv, bna_ctx *ctx, std::vector<std::shared_ptr<Req>> &rs)
{
    const Clock::time_point t0 = Clock::now();
    const int n = (int)rs.size();
    double work = 0.0;
    for (auto &r : rs) work += est_ms(*r);
    const bool big = srv->cfg.big_round_ms > 0.0 && work > srv->cfg.big_round_ms;
    if (big) {
        bna_set_chunking(ctx, srv->cfg.chunk_ms, gate_fn, srv);
        // leave one XCD's worth of CUs to the small rounds of other workers
        if (srv->grid_cap > 0) bna_set_grid_limit(ctx, srv->grid_cap);
    } else {
        std::lock_guard<std::mutex> lk(srv->gate_mu);
        srv->small_active++;
    }
    std::vector<bna_series *> ser(n, nullptr);
    std::vector<bna_opts> op(n);
    std::vector<bna_csr> dev(n);
    std::vector<int> status(n, BNA_OK);
    for (int q = 0; q < n; q++) {
        rs[q]->t_start = t0;
        op[q] = rs[q]->opts;
        status[q] = bna_series_create(ctx, rs[q]->x.data(), rs[q]->N, rs[q]->L, rs[q]->layout, 0, &ser[q]);
        // the host copy is on the device now; release it here, outside the server lock
        // (returning hundreds of MB to the OS takes milliseconds)
        std::vector<float>().swap(rs[q]->x);
    }
    const Clock::time_point t_up = Clock::now();
    bna_timings tm{};
    // the subjects that uploaded go through one batched build
    std::vector<int> ok;
    for (int q = 0; q < n; q++)
        if (status[q] == BNA_OK) ok.push_back(q);
    if (!ok.empty()) {
        std::vector<bna_series *> s2;
        std::vector<bna_opts> o2;
        for (int q : ok) {
            s2.push_back(ser[q]);
            o2.push_back(op[q]);
        }
        std::vector<bna_csr> d2(ok.size());
        std::vector<int> st2(ok.size(), BNA_OK);
        const int rc = bna_build_csr_batch(ctx, s2.data(), o2.data(), (int)ok.size(), d2.data(), st2.data());
        bna_get_timings(ctx, &tm);
        for (size_t t = 0; t < ok.size(); t++) {
            dev[ok[t]] = d2[t];
            status[ok[t]] = st2[t] != BNA_OK ? st2[t] : (rc != BNA_OK && !d2[t].R ? rc : BNA_OK);
        }
    }
    const Clock::time_point t_built = Clock::now();
    auto small_done = [&] {
        if (big) return;
        {
            std::lock_guard<std::mutex> lk(srv->gate_mu);
            srv->small_active--;
        }
        srv->gate_cv.notify_all();
    };
    for (int q = 0; q < n; q++) {
        Req &r = *rs[q];
        bna_result &res = r.res;
        memset(&res, 0, sizeof(res));
        res.status = status[q];
        res.batch_size = n;
        res.build = tm;
        res.upload_ms = ms_between(t0, t_up);
        res.build_ms = ms_between(t_up, t_built);
        const Clock::time_point tq = Clock::now();
        if (res.status == BNA_OK && r.want_degree) {
            res.degree = (float *)malloc(sizeof(float) * (size_t)r.N);
            res.status = res.degree ? bna_degree(ctx, &dev[q], res.degree) : BNA_ERR_NOMEM;
        }
        if (res.status == BNA_OK && r.want_ec) {
            res.ec = (double *)malloc(sizeof(double) * (size_t)r.N);
            res.status = res.ec ? bna_eigcent(ctx, &dev[q], 1e-6, 10000,
                                              dev[q].V ? BNA_EC_WEIGHTED_REF : BNA_EC_BINARY_REF, res.ec,
                                              &res.lambda, &res.ec_iters)
                                : BNA_ERR_NOMEM;
        }
        const Clock::time_point tm_done = Clock::now();
        res.maps_ms = ms_between(tq, tm_done);
        if (q == n - 1) small_done();  // the rest is host copies
        if (res.status == BNA_OK && r.device_result) {
            res.csr = dev[q];  // handed over; freed by bna_result_free
            memset(&dev[q], 0, sizeof(dev[q]));
        } else if (res.status == BNA_OK) {
            res.status = bna_csr_to_host(ctx, &dev[q], &res.csr);
        }
        const Clock::time_point t_dl = Clock::now();
        res.download_ms = ms_between(tm_done, t_dl);
        if (res.status != BNA_OK) {
            strncpy(res.error, bna_last_error(ctx), sizeof(res.error) - 1);
            bna_csr_free(&res.csr);
        }
        const Clock::time_point t1 = t_dl;
        res.queue_ms = ms_between(r.t_submit, r.t_start);
        res.run_ms = ms_between(r.t_start, t1);
        res.latency_ms = ms_between(r.t_submit, t1);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            r.done = true;  // the worker does not touch r.res from here on
            srv->st.completed++;
            if (res.status != BNA_OK) srv->st.failed++;
        }
        srv->cv_done.notify_all();
        // The request's device memory goes back to the context's pool after its
        // result is out: a free can wait behind another context's
        // synchronisation, and the caller should not wait with it.  Its footprint
        // is returned to the budget once the memory is.
        bna_csr_release(ctx, &dev[q]);
        if (ser[q]) bna_series_destroy(ctx, ser[q]);
        ser[q] = nullptr;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->reserved -= rs[q]->footprint;
        }
        srv->cv_work.notify_all();  // memory was released: more may be admitted
    }
    if (big) {
        bna_set_chunking(ctx, 0.0, nullptr, nullptr);
        bna_set_grid_limit(ctx, 0);
    }
    // keep the workspace from growing without bound across rounds
    if (bna_workspace_bytes(ctx) > (int64_t)(srv->cfg.hbm_budget / 4)) bna_trim(ctx);
}

static void worker(bna_server *srv, int w)
{
    bna_ctx *ctx = srv->ctxs[w];
    for (;;) {
        std::vector<std::shared_ptr<Req>> rs;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] {
                if (srv->stop) return true;
                if (srv->queue.empty()) return false;
                rs = admit(srv);
                return !rs.empty();
            });
            if (srv->stop && rs.empty()) return;
            srv->busy_workers++;
            srv->st.rounds++;
            srv->st.max_concurrent_rounds = std::max(srv->st.max_concurrent_rounds, (int64_t)srv->busy_workers);
            srv->st.max_reserved = std::max(srv->st.max_reserved, srv->reserved);
            srv->st.max_batch_seen = std::max(srv->st.max_batch_seen, (int64_t)rs.size());
        }
        run_round(srv, ctx, rs);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->busy_workers--;
        }
        srv->cv_work.notify_all();
    }
}

// One small build per precision, binary and weighted, then degree and
// centrality, on a worker's context before it takes requests: the first
// request a context serves then pays for no first-use set-up (kernel lookup,
// workspace growth, the centrality graph) that its solo time does not include.
static int warm_ctx(bna_ctx *ctx)
{
    const int N = 1000, L = 96;
    std::vector<float> x((size_t)N * L);
    (void)bna_synth_factor_model(x.data(), N, L, 8, 0.6, 0x5EA7ULL, 0);
    bna_series *s = nullptr;
    int rc = bna_series_create(ctx, x.data(), N, L, 0, 0, &s);
    if (rc) return rc;
    std::vector<float> deg(N);
    std::vector<double> ec(N);
    for (int prec = BNA_FP32; prec <= BNA_BF16X3 && rc == BNA_OK; prec++)
        for (int weighted = 0; weighted <= 1 && rc == BNA_OK; weighted++) {
            bna_opts o;
            bna_opts_default(&o);
            o.thr_type = BNA_THR_DENSITY;
            o.thr = 1.0;
            o.weighted = weighted;
            o.precision = prec;
            bna_csr g{};
            int st = BNA_OK;
            rc = bna_build_csr_batch(ctx, &s, &o, 1, &g, &st);
            if (rc == BNA_OK && prec == BNA_BF16X3) {
                double lam = 0.0;
                int it = 0;
                rc = bna_degree(ctx, &g, deg.data());
                if (rc == BNA_OK)
                    rc = bna_eigcent(ctx, &g, 1e-6, 10000, weighted ? BNA_EC_WEIGHTED_REF : BNA_EC_BINARY_REF,
                                     ec.data(), &lam, &it);
            }
            (void)bna_csr_release(ctx, &g);
        }
    (void)bna_series_destroy(ctx, s);
    return rc;
}

extern "C" {

void bna_server_config_default(bna_server_config *c)
{
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->device = 0;
    c->hbm_budget = 0;
    c->workers = 2;
    c->max_batch = 16;
    c->big_round_ms = 15.0;
    c->chunk_ms = 3.0;
    c->reserve_eighths = 1;
}

int bna_server_create(const bna_server_config *cfg, bna_server **out)
{
    if (!cfg || !out) return BNA_ERR_ARG;
    *out = nullptr;
    if (cfg->workers < 1 || cfg->workers > 8 || cfg->max_batch < 1 || cfg->reserve_eighths < 0 ||
        cfg->reserve_eighths > 7)
        return BNA_ERR_ARG;
    bna_server *srv = new (std::nothrow) bna_server();
    if (!srv) return BNA_ERR_NOMEM;
    srv->cfg = *cfg;
    for (int w = 0; w < cfg->workers; w++) {
        bna_ctx *c = nullptr;
        const int rc = bna_create(&c, cfg->device, 0);
        if (rc) {
            for (bna_ctx *x : srv->ctxs) bna_destroy(x);
            delete srv;
            return rc;
        }
        srv->ctxs.push_back(c);
    }
    if (srv->cfg.hbm_budget == 0) {
        int64_t fr = 0, tot = 0;
        if (bna_memory_info(srv->ctxs[0], &fr, &tot) == BNA_OK) srv->cfg.hbm_budget = (size_t)(0.8 * (double)fr);
    }
    {
        const int g = bna_grid_size(srv->ctxs[0]);
        if (g >= 16 && cfg->reserve_eighths > 0) srv->grid_cap = g - g / 8 * cfg->reserve_eighths;
    }
    for (bna_ctx *c : srv->ctxs) {
        (void)bna_reserve(c, cfg->reserve_per_worker);  // pool and staging
        const int rc = warm_ctx(c);
        if (rc) {
            for (bna_ctx *x : srv->ctxs) bna_destroy(x);
            delete srv;
            return rc;
        }
    }
    for (int w = 0; w < cfg->workers; w++) srv->workers.emplace_back(worker, srv, w);
    *out = srv;
    return BNA_OK;
}

int bna_server_destroy(bna_server *srv)
{
    if (!srv) return BNA_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        for (auto &r : srv->queue) {  // never admitted: nothing reserved
            r->res.status = BNA_ERR_CANCELLED;
            r->done = true;
        }
        srv->queue.clear();
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers) t.join();
    for (auto &kv : srv->all)
        if (kv.second->done) bna_result_free(&kv.second->res);
    for (bna_ctx *c : srv->ctxs) bna_destroy(c);
    delete srv;
    return BNA_OK;
}

int bna_submit(bna_server *srv, const bna_request *req, int64_t *id)
{
    if (!srv || !req || !id || !req->x || req->N < 2 || req->L < 2) return BNA_ERR_ARG;
    if (req->layout != BNA_TIME_MAJOR && req->layout != BNA_VOXEL_MAJOR) return BNA_ERR_ARG;
    if (req->opts.precision < BNA_FP32 || req->opts.precision > BNA_BF16X3) return BNA_ERR_ARG;
    auto r = std::make_shared<Req>();
    try {
        r->x.assign(req->x, req->x + (size_t)req->N * req->L);
    } catch (...) {
        return BNA_ERR_NOMEM;
    }
    r->N = req->N;
    r->L = req->L;
    r->layout = req->layout;
    r->opts = req->opts;
    r->want_degree = req->want_degree;
    r->want_ec = req->want_ec;
    r->device_result = req->device_result;
    r->footprint = bna_request_footprint(req->N, req->L, &req->opts);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return BNA_ERR_CANCELLED;
        r->id = srv->next_id++;
        srv->queue.push_back(r);
        srv->all[r->id] = r;
        srv->st.submitted++;
        *id = r->id;
    }
    srv->cv_work.notify_all();
    return BNA_OK;
}

static int collect(bna_server *srv, int64_t id, bna_result *out, bool wait)
{
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) return BNA_ERR_NOT_FOUND;
    std::shared_ptr<Req> r = it->second;
    if (!r->done) {
        if (!wait) return BNA_ERR_BUSY;
        srv->cv_done.wait(lk, [&] { return r->done; });
    }
    *out = r->res;
    memset(&r->res, 0, sizeof(r->res));
    srv->all.erase(id);
    return out->status;
}

int bna_poll(bna_server *srv, int64_t id, bna_result *out)
{
    if (!srv || !out) return BNA_ERR_ARG;
    memset(out, 0, sizeof(*out));
    return collect(srv, id, out, false);
}

int bna_wait(bna_server *srv, int64_t id, bna_result *out)
{
    if (!srv || !out) return BNA_ERR_ARG;
    memset(out, 0, sizeof(*out));
    return collect(srv, id, out, true);
}

int bna_cancel(bna_server *srv, int64_t id)
{
    if (!srv) return BNA_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) return BNA_ERR_NOT_FOUND;
    for (auto q = srv->queue.begin(); q != srv->queue.end(); ++q)
        if ((*q)->id == id) {
            (*q)->res.status = BNA_ERR_CANCELLED;
            (*q)->res.latency_ms = ms_between((*q)->t_submit, Clock::now());
            srv->queue.erase(q);
            it->second->footprint = 0;
            it->second->done = true;
            it->second->x.clear();
            srv->st.cancelled++;
            srv->cv_done.notify_all();
            return BNA_OK;
        }
    return it->second->done ? BNA_OK : BNA_ERR_BUSY;  // running requests finish
}

int bna_server_get_stats(bna_server *srv, bna_server_stats *out)
{
    if (!srv || !out) return BNA_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->st;
    out->budget = (int64_t)srv->cfg.hbm_budget;
    out->queued = (int64_t)srv->queue.size();
    return BNA_OK;
}

void bna_result_free(bna_result *r)
{
    if (!r) return;
    bna_csr_free(&r->csr);
    free(r->degree);
    free(r->ec);
    r->degree = nullptr;
    r->ec = nullptr;
}

}  // extern "C"
