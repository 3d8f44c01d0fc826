// serve.hip -- continuous-batching SpGEMM server (tsg_server_*).
//
// A request is C = A^2 or C = A*A^T of a matrix registered once: its CSR,
// its tiles (with CSC and row indices) and those of A^T stay resident in
// HBM, and registration sizes each product exactly by running it once
// (the symbolic steps give C's tiles and nonzeros; the chunk plan gives
// the tile-row pieces of products whose C exceeds the int32 offsets of
// the tile format).  That exact size is the request's admission estimate.
//
// `slots` worker threads, each with its own context (HIP stream and
// allocator), pull work whenever they are idle: the first queued request
// whose bytes fit the budget left (budget minus resident operands minus
// admitted requests), and -- when it is small -- every later queued small
// request that also fits, up to max_batch, fused into one block-diagonal
// product (tsg_tiles_pack of the resident tiles, one pipeline for all of
// them).  Requests arriving while others run are therefore admitted as
// soon as a slot and memory free up, and their work overlaps on the
// device through the slots' streams.  A request that does not fit is
// passed over by smaller ones until it has waited starve_ms; then nothing
// behind it is admitted before it.  Each result (C tiles) stays in HBM
// until released and is bit-identical to tsg_spgemm on the same operands.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tsg_internal.h"

namespace {

double clock_ms() {
    static const auto t0 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Device bytes of a C tile matrix plus the pipeline's work buffer.
long long c_bytes(long long tilem, long long numtile, long long nnz, int s1w) {
    tsg_dtiles c;
    memset(&c, 0, sizeof(c));
    c.tilem = (int)tilem;
    c.numtile = (int)numtile;
    c.nnz = (int)nnz;
    const long long arrays = 4 * (tilem + 1) + 12 * (numtile + 1) + 48 * numtile + 32 + 9 * (nnz + 1);
    return arrays + (long long)tsg_spgemm_work_bytes(&c, nullptr, s1w);
}

// Device bytes a tile matrix allocated on ctx holds.
long long held_bytes(const tsg_ctx *ctx, const tsg_dtiles &t) {
    const void *arr[] = {t.tile_ptr, t.tile_colidx, t.tile_rowidx, t.tile_nnz, t.csr_ptr, t.csr_col, t.val,
                         t.mask, t.csc_ptr, t.csc_rowidx, t.csc_tileid, t.rowx_ptr, t.rowx_col, t.rowx_val};
    long long b = 0;
    for (const void *p : arr)
        if (p) b += (long long)tsg_alloc_size(ctx, p);
    return b;
}

long long tile_bytes(const tsg_dtiles &t) {
    long long b = 4LL * (t.tilem + 1) + 12LL * (t.numtile + 1) + 48LL * t.numtile + 9LL * t.nnz;
    if (t.csc_ptr) b += 4LL * (t.tilen + 1) + 8LL * t.numtile;
    if (t.rowx_ptr) b += 4LL * (t.m + 1) + 12LL * t.nnz;
    return b;
}

struct OpSize {
    bool ok = false;
    int status = TSG_OK;
    std::vector<int> bounds;  // tile-row chunk bounds
    long long numtileC = 0, nnzC = 0, bytes = 0;
};

struct Matrix {
    tsg_dcsr a, at;
    tsg_dtiles ta, tat;
    OpSize size[2];
    long long resident = 0;
};

struct Fused;

struct Request {
    unsigned long long id;
    int handle, op;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    double t_gpu = 0;  // device time of its pipeline runs
    int status = TSG_OK;
    long long bytes = 0;
    bool done = false, released = false, auto_release = false;
    int worker = -1;
    int fused_with = 1;
    // result: own chunks, or a share of a fused product
    std::vector<tsg_dtiles> chunks;
    std::vector<int> chunk_r0;
    std::shared_ptr<Fused> fused;
    int f_tr = 0, f_tc = 0;  // tile-row / tile-column base inside the fused C
};

struct Fused {
    tsg_dtiles C;
    int worker;
    int live;  // requests still holding it
};

}  // namespace

struct tsg_server {
    tsg_server_config cfg;
    int device;
    tsg_ctx *main = nullptr;  // registration and result downloads
    std::mutex main_mu;
    std::vector<tsg_ctx *> wctx;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::vector<std::unique_ptr<Matrix>> mats;
    std::deque<unsigned long long> queue;
    std::unordered_map<unsigned long long, std::unique_ptr<Request>> reqs;
    unsigned long long next_id = 1;
    long long resident = 0, reserved = 0, peak_reserved = 0;
    long long fused_launches = 0, single_launches = 0;
    // device memory to free, per worker (allocators are per context)
    std::vector<std::vector<tsg_dtiles>> pending_free;
    // free blocks each worker's allocator holds, and trim requests: cached
    // blocks count against the budget, so that resident + reserved + cached
    // bounds the device memory the server uses
    std::vector<long long> wcached;
    std::vector<char> trim_req;
    long long peak_footprint = 0;
    // every context allocates from one arena of the budget's size, reserved
    // at creation: trimming a cache returns blocks to it, never to the driver
    tsg_arena *arena = nullptr;
};

extern "C" void tsg_server_config_default(tsg_server_config *c) {
    memset(c, 0, sizeof(*c));
    c->budget_bytes = 64LL << 30;
    c->slots = 4;
    c->max_batch = 64;
    c->small_bytes = 256LL << 20;
    c->chunk_products = 1500000000LL;
    c->starve_ms = 200.0;
    tsg_tune_default(&c->tune);
}

static void worker_main(tsg_server *s, int w);

extern "C" int tsg_server_create(int device, const tsg_server_config *cfg, tsg_server **out) {
    *out = nullptr;
    tsg_server *s = new tsg_server();
    if (cfg) s->cfg = *cfg; else tsg_server_config_default(&s->cfg);
    if (s->cfg.slots < 1) s->cfg.slots = 1;
    if (s->cfg.max_batch < 1) s->cfg.max_batch = 1;
    if (s->cfg.chunk_products < 1) s->cfg.chunk_products = 1500000000LL;
    s->device = device;
    int st = tsg_arena_create(device, (size_t)std::max(1LL, s->cfg.budget_bytes), &s->arena);
    if (st) { delete s; return st; }
    st = tsg_ctx_create(device, &s->main);
    if (st) {
        tsg_arena_destroy(s->arena);
        delete s;
        return st;
    }
    tsg_ctx_set_arena(s->main, s->arena);
    s->wctx.resize(s->cfg.slots, nullptr);
    s->pending_free.resize(s->cfg.slots);
    s->wcached.assign(s->cfg.slots, 0);
    s->trim_req.assign(s->cfg.slots, 0);
    for (int w = 0; w < s->cfg.slots; w++) {
        st = tsg_ctx_create(device, &s->wctx[w]);
        if (st) break;
        tsg_ctx_set_arena(s->wctx[w], s->arena);
    }
    if (st) {
        for (auto *c : s->wctx) if (c) tsg_ctx_destroy(c);
        tsg_ctx_destroy(s->main);
        tsg_arena_destroy(s->arena);
        delete s;
        return st;
    }
    for (int w = 0; w < s->cfg.slots; w++) s->workers.emplace_back(worker_main, s, w);
    *out = s;
    return TSG_OK;
}

// Sizes A*B exactly from the symbolic steps: every chunk runs steps 1-2 and
// the numeric step's plan (row classes, heavy-row work list), not the
// numeric step itself.  C's tiles and nonzeros per chunk, its element
// segments and heavy-row scratch give the request's admission bytes.
static int size_product(tsg_server *s, tsg_ctx *ctx, const tsg_dtiles &A, const tsg_dtiles &B, OpSize &os) {
    int nb_max = 1 << 16;
    std::vector<int> bounds(nb_max + 1);
    int n = 0;
    TSG_TRY(tsg_spgemm_chunk_plan(ctx, &A, &B, s->cfg.chunk_products, nb_max, bounds.data(), &n, nullptr));
    bounds.resize(n + 1);
    os.bounds = bounds;
    os.numtileC = os.nnzC = os.bytes = 0;
    long long transient = 0;
    for (int c = 0; c < n; c++) {
        tsg_dtiles C;
        tsg_stats st;
        TSG_TRY(tsg_spgemm_tile_rows_ex(ctx, &A, &B, bounds[c], bounds[c + 1], &s->cfg.tune, &C, &st, true));
        os.numtileC += st.numtileC;
        os.nnzC += st.nnzC;
        os.bytes += c_bytes(C.tilem, st.numtileC, st.nnzC, tsg_s1_words(&B, s->cfg.tune));
        transient = std::max(transient, st.heavy_bytes);  // freed after each chunk
        tsg_tiles_free(ctx, &C);
    }
    os.bytes += transient + os.bytes / 32 + (1 << 20);  // + allocator rounding and small scratch
    os.ok = true;
    return TSG_OK;
}

extern "C" int tsg_server_register(tsg_server *s, const tsg_hcsr *h, int *handle) {
    if (!s || !h || !handle) { tsg_set_error("tsg_server_register: null argument"); return TSG_ERR_ARG; }
    std::lock_guard<std::mutex> g(s->main_mu);
    tsg_ctx *ctx = s->main;
    auto M = std::make_unique<Matrix>();
    memset(&M->a, 0, sizeof(M->a));
    memset(&M->at, 0, sizeof(M->at));
    memset(&M->ta, 0, sizeof(M->ta));
    memset(&M->tat, 0, sizeof(M->tat));
    auto fail = [&](int st) {
        tsg_tiles_free(ctx, &M->tat);
        tsg_tiles_free(ctx, &M->ta);
        tsg_dcsr_free(ctx, &M->at);
        tsg_dcsr_free(ctx, &M->a);
        return st;
    };
    int st = tsg_dcsr_upload(ctx, h, &M->a);
    if (!st) st = tsg_validate_csr(ctx, &M->a);
    if (!st) st = tsg_csr2tile(ctx, &M->a, &M->ta);
    if (!st) st = tsg_tiles_build_csc(ctx, &M->ta);
    if (!st) st = tsg_tiles_build_rowindex(ctx, &M->ta);
    if (!st) st = tsg_dcsr_transpose(ctx, &M->a, &M->at);
    if (!st) st = tsg_csr2tile(ctx, &M->at, &M->tat);
    if (!st) st = tsg_tiles_build_csc(ctx, &M->tat);
    if (!st) st = tsg_tiles_build_rowindex(ctx, &M->tat);
    if (st) return fail(st);
    M->resident = 12LL * (M->a.m + 1 + M->a.n + 1) + 24LL * M->a.nnz + tile_bytes(M->ta) + tile_bytes(M->tat);
    // exact sizes of A^2 (square only) and A*A^T
    if (M->a.m == M->a.n) {
        st = size_product(s, ctx, M->ta, M->ta, M->size[TSG_OP_A2]);
        if (st) M->size[TSG_OP_A2].status = st;
    } else {
        M->size[TSG_OP_A2].status = TSG_ERR_ARG;
    }
    st = size_product(s, ctx, M->ta, M->tat, M->size[TSG_OP_AAT]);
    if (st) M->size[TSG_OP_AAT].status = st;
    tsg_ctx_sync(ctx);
    tsg_ctx_trim(ctx);
    std::lock_guard<std::mutex> g2(s->mu);
    if (s->resident + M->resident > s->cfg.budget_bytes) {
        tsg_set_error("tsg_server_register: resident operands (%lld bytes) exceed the budget", s->resident + M->resident);
        return fail(TSG_ERR_BUDGET);
    }
    s->resident += M->resident;
    s->mats.push_back(std::move(M));
    *handle = (int)s->mats.size() - 1;
    return TSG_OK;
}

extern "C" int tsg_server_size(tsg_server *s, int handle, int op, long long *numtileC, long long *nnzC,
                               long long *bytes, int *chunks) {
    std::lock_guard<std::mutex> g(s->mu);
    if (handle < 0 || handle >= (int)s->mats.size() || (op != TSG_OP_A2 && op != TSG_OP_AAT)) {
        tsg_set_error("tsg_server_size: bad handle or op");
        return TSG_ERR_ARG;
    }
    const OpSize &o = s->mats[handle]->size[op];
    if (!o.ok) { tsg_set_error("tsg_server_size: product not available (status %d)", o.status); return o.status ? o.status : TSG_ERR_ARG; }
    if (numtileC) *numtileC = o.numtileC;
    if (nnzC) *nnzC = o.nnzC;
    if (bytes) *bytes = o.bytes;
    if (chunks) *chunks = (int)o.bounds.size() - 1;
    return TSG_OK;
}

static int submit_locked(tsg_server *s, int handle, int op, bool auto_release, unsigned long long *id) {
    if (handle < 0 || handle >= (int)s->mats.size() || (op != TSG_OP_A2 && op != TSG_OP_AAT)) {
        tsg_set_error("tsg_server_submit: bad handle or op");
        return TSG_ERR_ARG;
    }
    const OpSize &o = s->mats[handle]->size[op];
    if (!o.ok) { tsg_set_error("tsg_server_submit: product not available (status %d)", o.status); return o.status ? o.status : TSG_ERR_ARG; }
    if (s->resident + o.bytes > s->cfg.budget_bytes) {
        tsg_set_error("tsg_server_submit: request needs %lld bytes, budget left after residents %lld", o.bytes,
                      s->cfg.budget_bytes - s->resident);
        return TSG_ERR_BUDGET;
    }
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->handle = handle;
    r->op = op;
    r->bytes = o.bytes;
    r->auto_release = auto_release;
    r->t_arrival = clock_ms();
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->cv_work.notify_one();
    return TSG_OK;
}

extern "C" int tsg_server_submit(tsg_server *s, int handle, int op, unsigned long long *id) {
    std::lock_guard<std::mutex> g(s->mu);
    return submit_locked(s, handle, op, false, id);
}

// Under s->mu: bytes cached by the workers other than w.
static long long others_cached(const tsg_server *s, int w) {
    long long c = 0;
    for (int v = 0; v < (int)s->wcached.size(); v++)
        if (v != w) c += s->wcached[v];
    return c;
}

// Caches with a trim pending are about to be freed before their worker's
// next job allocates anything.
static void note_footprint(tsg_server *s) {
    long long c = 0;
    for (int v = 0; v < (int)s->wcached.size(); v++)
        if (!s->trim_req[v]) c += s->wcached[v];
    s->peak_footprint = std::max(s->peak_footprint, s->resident + s->reserved + c);
}

// Picks the next job for worker w under s->mu: returns the admitted
// requests.  Worker w's own cached blocks are room for its job (it trims
// them if the job does not fit beside them); the other workers' count as
// used.  When only those caches keep the queue head out, their workers are
// asked to trim.
static std::vector<Request *> pick(tsg_server *s, int w) {
    std::vector<Request *> job;
    const long long free_budget = s->cfg.budget_bytes - s->resident - s->reserved;
    const long long avail = free_budget - others_cached(s, w);
    const double now = clock_ms();
    if (!s->queue.empty()) {
        const Request *head = s->reqs[s->queue[0]].get();
        if (head->bytes > avail && head->bytes <= free_budget)
            for (int v = 0; v < (int)s->wcached.size(); v++)
                if (v != w && s->wcached[v] > 0) s->trim_req[v] = 1;
    }
    size_t qi = 0;
    for (; qi < s->queue.size(); qi++) {
        Request *r = s->reqs[s->queue[qi]].get();
        if (r->bytes <= avail) break;
        if (qi == 0 && now - r->t_arrival > s->cfg.starve_ms) return job;  // reserve for the head
    }
    if (qi == s->queue.size()) return job;
    Request *first = s->reqs[s->queue[qi]].get();
    job.push_back(first);
    long long used = first->bytes;
    if (first->bytes <= s->cfg.small_bytes) {
        for (size_t j = qi + 1; j < s->queue.size() && (int)job.size() < s->cfg.max_batch; j++) {
            Request *r = s->reqs[s->queue[j]].get();
            if (r->bytes <= s->cfg.small_bytes && used + r->bytes <= avail) {
                job.push_back(r);
                used += r->bytes;
            }
        }
    }
    for (Request *r : job) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r->id));
        r->t_admit = now;
        s->reserved += r->bytes;
    }
    s->peak_reserved = std::max(s->peak_reserved, s->reserved);
    // the worker's cache beside the job must fit what is left; else it is
    // trimmed before the job runs (other workers count it until then)
    if (s->wcached[w] > s->cfg.budget_bytes - s->resident - s->reserved - others_cached(s, w)) s->trim_req[w] = 1;
    note_footprint(s);
    return job;
}

static void free_request_results(tsg_server *s, Request *r) {
    // under s->mu: hand device memory to its worker
    // (the blocks reach the worker's cache when it frees them: counted there now)
    for (auto &c : r->chunks) s->pending_free[r->worker].push_back(c);
    r->chunks.clear();
    if (r->fused) {
        if (--r->fused->live == 0) s->pending_free[r->fused->worker].push_back(r->fused->C);
        r->fused.reset();
    }
    s->wcached[r->worker] += r->bytes;
    s->reserved -= r->bytes;
    r->bytes = 0;
}

static int run_single(tsg_server *s, tsg_ctx *ctx, Request *r) {
    Matrix &M = *s->mats[r->handle];
    const tsg_dtiles &B = r->op == TSG_OP_A2 ? M.ta : M.tat;
    const OpSize &o = M.size[r->op];
    for (size_t c = 0; c + 1 < o.bounds.size(); c++) {
        tsg_dtiles C;
        tsg_stats st;
        TSG_TRY(tsg_spgemm_tile_rows(ctx, &M.ta, &B, o.bounds[c], o.bounds[c + 1], &s->cfg.tune, &C, &st));
        r->t_gpu += st.t_total_ms;
        r->chunks.push_back(C);
        r->chunk_r0.push_back(o.bounds[c]);
    }
    return tsg_ctx_sync(ctx);
}

static int run_fused(tsg_server *s, tsg_ctx *ctx, std::vector<Request *> &job, std::shared_ptr<Fused> &f) {
    const int n = (int)job.size();
    std::vector<const tsg_dtiles *> pa(n), pb(n);
    for (int i = 0; i < n; i++) {
        Matrix &M = *s->mats[job[i]->handle];
        pa[i] = &M.ta;
        pb[i] = job[i]->op == TSG_OP_A2 ? &M.ta : &M.tat;
    }
    tsg_dtiles A, B;
    TSG_TRY(tsg_tiles_pack(ctx, pa.data(), n, &A));
    int st = tsg_tiles_pack(ctx, pb.data(), n, &B);
    if (st) { tsg_tiles_free(ctx, &A); return st; }
    f = std::make_shared<Fused>();
    f->live = n;
    tsg_stats ss;
    st = tsg_spgemm(ctx, &A, &B, &s->cfg.tune, &f->C, &ss);
    if (!st)
        for (Request *r : job) r->t_gpu = ss.t_total_ms;
    if (!st) st = tsg_ctx_sync(ctx);
    tsg_tiles_free(ctx, &A);
    tsg_tiles_free(ctx, &B);
    if (st) return st;
    int tr = 0, tc = 0;
    for (int i = 0; i < n; i++) {
        job[i]->f_tr = tr;
        job[i]->f_tc = tc;
        tr += pa[i]->tilem;
        tc += pb[i]->tilen;
    }
    return TSG_OK;
}

static void worker_main(tsg_server *s, int w) {
    tsg_ctx *ctx = s->wctx[w];
    hipSetDevice(s->device);
    for (;;) {
        std::vector<Request *> job;
        std::vector<tsg_dtiles> to_free;
        bool trim = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                to_free.swap(s->pending_free[w]);
                if (s->trim_req[w]) {
                    trim = true;
                    s->trim_req[w] = 0;
                }
                if (trim || !to_free.empty() || s->stop) break;
                job = pick(s, w);
                if (!job.empty()) {
                    trim = s->trim_req[w];  // the job needs this worker's cache trimmed first
                    break;
                }
                if (std::find(s->trim_req.begin(), s->trim_req.end(), 1) != s->trim_req.end())
                    s->cv_work.notify_all();  // other workers were asked to trim
                s->cv_work.wait_for(lk, std::chrono::milliseconds(20));
            }
            if (job.empty() && to_free.empty() && !trim && s->stop) return;
        }
        for (auto &t : to_free) tsg_tiles_free(ctx, &t);
        if (trim) {
            tsg_ctx_trim(ctx);
            std::lock_guard<std::mutex> g(s->mu);
            s->wcached[w] = tsg_ctx_bytes_cached(ctx);
            s->trim_req[w] = 0;
        }
        if (job.empty()) {
            std::lock_guard<std::mutex> g(s->mu);
            s->wcached[w] = tsg_ctx_bytes_cached(ctx);
            note_footprint(s);
            s->cv_work.notify_all();  // memory came back
            continue;
        }
        int st;
        std::shared_ptr<Fused> f;
        if (job.size() == 1) {
            st = run_single(s, ctx, job[0]);
        } else {
            st = run_fused(s, ctx, job, f);
            if (f) f->worker = w;
        }
        long long held = -1;  // a single run's result: its transient blocks are now cached
        if (!st && job.size() == 1) {
            held = 0;
            for (auto &c : job[0]->chunks) held += held_bytes(ctx, c);
        }
        std::unique_lock<std::mutex> g(s->mu);
        if (held >= 0 && held < job[0]->bytes) {
            s->reserved -= job[0]->bytes - held;
            job[0]->bytes = held;
        }
        s->wcached[w] = tsg_ctx_bytes_cached(ctx);
        if (s->resident + s->reserved + others_cached(s, w) + s->wcached[w] > s->cfg.budget_bytes) {
            // the cache beside what is held would exceed the budget
            g.unlock();
            tsg_ctx_trim(ctx);
            g.lock();
            s->wcached[w] = tsg_ctx_bytes_cached(ctx);
        }
        note_footprint(s);
        const double now = clock_ms();
        if (job.size() == 1) s->single_launches++; else s->fused_launches++;
        for (Request *r : job) {
            r->worker = w;
            r->t_done = now;
            r->status = st;
            r->fused_with = (int)job.size();
            r->done = true;
            if (st) {  // no result: give back its memory now
                for (auto &c : r->chunks) s->pending_free[w].push_back(c);
                r->chunks.clear();
                s->wcached[w] += r->bytes;
                s->reserved -= r->bytes;
                r->bytes = 0;
                r->released = true;
            } else {
                if (job.size() > 1) r->fused = f;
                if (r->auto_release) {
                    free_request_results(s, r);
                    r->released = true;
                }
            }
        }
        if (st && f) s->pending_free[w].push_back(f->C);  // partial C of a failed fused run
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

extern "C" int tsg_server_wait(tsg_server *s, unsigned long long id, double timeout_ms, tsg_request_info *info) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { tsg_set_error("tsg_server_wait: unknown request %llu", id); return TSG_ERR_ARG; }
    Request *r = it->second.get();
    const auto until = std::chrono::steady_clock::now() +
                       std::chrono::microseconds((long long)(timeout_ms > 0 ? timeout_ms * 1000 : 1e12));
    while (!r->done) {
        if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout && !r->done) {
            tsg_set_error("tsg_server_wait: request %llu not done within %.0f ms", id, timeout_ms);
            return TSG_ERR_STATE;
        }
    }
    if (info) {
        memset(info, 0, sizeof(*info));
        const OpSize &o = s->mats[r->handle]->size[r->op];
        info->status = r->status;
        info->handle = r->handle;
        info->op = r->op;
        info->numtileC = o.numtileC;
        info->nnzC = o.nnzC;
        info->chunks = (int)o.bounds.size() - 1;
        info->fused_with = r->fused_with;
        info->bytes = o.bytes;
        info->t_arrival_ms = r->t_arrival;
        info->t_admit_ms = r->t_admit;
        info->t_done_ms = r->t_done;
        info->t_gpu_ms = r->t_gpu;
    }
    return r->status;
}

extern "C" int tsg_server_release(tsg_server *s, unsigned long long id) {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { tsg_set_error("tsg_server_release: unknown request %llu", id); return TSG_ERR_ARG; }
    Request *r = it->second.get();
    if (!r->done) { tsg_set_error("tsg_server_release: request %llu still running", id); return TSG_ERR_STATE; }
    if (!r->released) free_request_results(s, r);
    s->reqs.erase(it);
    s->cv_work.notify_all();
    return TSG_OK;
}

// The C views of a finished, unreleased request: its chunks (first tile-row
// of each in r0s) or its share of a fused product (one view, tile columns
// relative to jb).
static int result_views(tsg_server *s, unsigned long long id, const char *who, std::vector<tsg_dtiles> &views,
                        std::vector<int> &jb, std::vector<int> &r0s, int *m_out, int *n_out) {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { tsg_set_error("%s: unknown request", who); return TSG_ERR_ARG; }
    Request *r = it->second.get();
    if (!r->done || r->released || r->status) {
        tsg_set_error("%s: request %llu has no result", who, id);
        return TSG_ERR_STATE;
    }
    const Matrix &M = *s->mats[r->handle];
    const int m = M.a.m;
    const int n = r->op == TSG_OP_A2 ? M.a.n : M.a.m;
    *m_out = m;
    *n_out = n;
    if (r->fused) {
        tsg_dtiles v = r->fused->C;
        v.tile_ptr = r->fused->C.tile_ptr + r->f_tr;
        v.m = m;
        v.tilem = (m + 15) / 16;
        v.n = n;
        v.tilen = (n + 15) / 16;
        views.push_back(v);
        jb.push_back(r->f_tc);
        r0s.push_back(0);
    } else {
        for (size_t c = 0; c < r->chunks.size(); c++) {
            views.push_back(r->chunks[c]);
            jb.push_back(0);
            r0s.push_back(r->chunk_r0[c]);
        }
    }
    return TSG_OK;
}

extern "C" int tsg_server_result_checksum(tsg_server *s, unsigned long long id, tsg_csum *out) {
    memset(out, 0, sizeof(*out));
    std::vector<tsg_dtiles> views;
    std::vector<int> jb, r0s;
    int m = 0, n = 0;
    TSG_TRY(result_views(s, id, "tsg_server_result_checksum", views, jb, r0s, &m, &n));
    std::lock_guard<std::mutex> g(s->main_mu);
    for (size_t i = 0; i < views.size(); i++) {
        tsg_csum c;
        TSG_TRY(tsg_tiles_checksum(s->main, &views[i], 16LL * r0s[i], jb[i], &c));
        out->numtile += c.numtile;
        out->nnz += c.nnz;
        out->tiles += c.tiles;
        out->entries += c.entries;
        out->entries_t += c.entries_t;
    }
    return TSG_OK;
}

extern "C" int tsg_server_result_rows(tsg_server *s, unsigned long long id, const int *rows, int nrows,
                                      tsg_hcsr *out) {
    memset(out, 0, sizeof(*out));
    if (nrows < 0 || (nrows > 0 && !rows)) { tsg_set_error("tsg_server_result_rows: bad row list"); return TSG_ERR_ARG; }
    std::vector<tsg_dtiles> views;
    std::vector<int> jb, r0s;
    int m = 0, n = 0;
    TSG_TRY(result_views(s, id, "tsg_server_result_rows", views, jb, r0s, &m, &n));
    for (int x = 0; x < nrows; x++)
        if (rows[x] < 0 || rows[x] >= m) {
            tsg_set_error("tsg_server_result_rows: row %d outside [0, %d)", rows[x], m);
            return TSG_ERR_ARG;
        }
    // rows of each view, then reassembled in the caller's order
    std::vector<std::vector<int>> local(views.size()), pos(views.size());
    for (int x = 0; x < nrows; x++) {
        size_t v = views.size() - 1;
        while (v > 0 && 16 * r0s[v] > rows[x]) v--;
        local[v].push_back(rows[x] - 16 * r0s[v]);
        pos[v].push_back(x);
    }
    std::vector<tsg_hcsr> parts(views.size());
    std::vector<int> len(nrows, 0), from_part(nrows, 0), from_row(nrows, 0);
    int st = TSG_OK;
    {
        std::lock_guard<std::mutex> g(s->main_mu);
        for (size_t v = 0; v < views.size() && st == TSG_OK; v++) {
            memset(&parts[v], 0, sizeof(parts[v]));
            if (local[v].empty()) continue;
            st = tsg_tiles_rows(s->main, &views[v], jb[v], local[v].data(), (int)local[v].size(), &parts[v]);
        }
        tsg_ctx_trim(s->main);
    }
    long long nnz = 0;
    if (st == TSG_OK) {
        for (size_t v = 0; v < views.size(); v++)
            for (size_t k = 0; k < pos[v].size(); k++) {
                const int x = pos[v][k];
                len[x] = parts[v].rowptr[k + 1] - parts[v].rowptr[k];
                from_part[x] = (int)v;
                from_row[x] = (int)k;
                nnz += len[x];
            }
        out->m = nrows;
        out->n = n;
        out->nnz = (int)nnz;
        out->rowptr = (int *)malloc(4 * ((size_t)nrows + 1));
        out->colidx = (int *)malloc(4 * (size_t)(nnz ? nnz : 1));
        out->val = (double *)malloc(8 * (size_t)(nnz ? nnz : 1));
        if (!out->rowptr || !out->colidx || !out->val) {
            tsg_set_error("tsg_server_result_rows: host allocation failed");
            st = TSG_ERR_NOMEM;
        }
    }
    if (st == TSG_OK) {
        long long e = 0;
        for (int x = 0; x < nrows; x++) {
            out->rowptr[x] = (int)e;
            const tsg_hcsr &p = parts[from_part[x]];
            const int s0 = p.rowptr[from_row[x]];
            memcpy(out->colidx + e, p.colidx + s0, 4 * (size_t)len[x]);
            memcpy(out->val + e, p.val + s0, 8 * (size_t)len[x]);
            e += len[x];
        }
        out->rowptr[nrows] = (int)e;
    }
    for (auto &p : parts) tsg_hcsr_free(&p);
    if (st != TSG_OK) tsg_hcsr_free(out);
    return st;
}

// C of a finished, unreleased request as a host CSR (rows of every chunk
// concatenated).
extern "C" int tsg_server_result_csr(tsg_server *s, unsigned long long id, tsg_hcsr *out) {
    memset(out, 0, sizeof(*out));
    std::vector<tsg_dtiles> views;
    std::vector<int> jb, r0s;
    int m = 0, n = 0;
    TSG_TRY(result_views(s, id, "tsg_server_result_csr", views, jb, r0s, &m, &n));
    std::lock_guard<std::mutex> g(s->main_mu);
    tsg_ctx *ctx = s->main;
    std::vector<tsg_hcsr> parts(views.size());
    long long nnz = 0;
    for (size_t i = 0; i < views.size(); i++) {
        tsg_dcsr d;
        TSG_TRY(tsg_tile2csr_view(ctx, &views[i], jb[i], &d));
        int st = tsg_dcsr_download(ctx, &d, &parts[i]);
        tsg_dcsr_free(ctx, &d);
        TSG_TRY(st);
        nnz += parts[i].nnz;
    }
    tsg_ctx_trim(ctx);  // the CSR staging goes back to the arena
    out->m = m;
    out->n = n;
    out->nnz = (int)nnz;
    out->rowptr = (int *)malloc(4 * ((size_t)m + 1));
    out->colidx = (int *)malloc(4 * (size_t)(nnz ? nnz : 1));
    out->val = (double *)malloc(8 * (size_t)(nnz ? nnz : 1));
    if (!out->rowptr || !out->colidx || !out->val) {
        for (auto &p : parts) tsg_hcsr_free(&p);
        tsg_hcsr_free(out);
        tsg_set_error("tsg_server_result_csr: host allocation failed");
        return TSG_ERR_NOMEM;
    }
    int row = 0;
    long long e = 0;
    for (auto &p : parts) {
        for (int i = 0; i < p.m && row < m; i++) out->rowptr[row++] = (int)(e + p.rowptr[i]);
        memcpy(out->colidx + e, p.colidx, 4 * (size_t)p.nnz);
        memcpy(out->val + e, p.val, 8 * (size_t)p.nnz);
        e += p.nnz;
        tsg_hcsr_free(&p);
    }
    out->rowptr[m] = (int)e;
    return TSG_OK;
}

extern "C" int tsg_server_stats(tsg_server *s, tsg_server_stats_t *out) {
    std::lock_guard<std::mutex> g(s->mu);
    memset(out, 0, sizeof(*out));
    out->resident_bytes = s->resident;
    out->reserved_bytes = s->reserved;
    out->peak_reserved_bytes = s->peak_reserved;
    out->queued = (int)s->queue.size();
    out->fused_launches = s->fused_launches;
    out->single_launches = s->single_launches;
    for (long long c : s->wcached) out->cached_bytes += c;
    out->peak_footprint_bytes = s->peak_footprint;
    return TSG_OK;
}

// Open-loop load generator: request i (handles[i], ops[i]) is submitted at
// arrival_ms[i] after the call starts; returns when every request is done.
// keep = 0 releases each result as it completes (its memory returns to the
// budget); keep = 1 leaves them for tsg_server_result_csr / release.
extern "C" int tsg_server_replay(tsg_server *s, int n, const int *handles, const int *ops, const double *arrival_ms,
                                 int keep, unsigned long long *ids, tsg_request_info *infos) {
    return tsg_server_replay_progress(s, n, handles, ops, arrival_ms, keep, ids, infos, nullptr, nullptr);
}

extern "C" int tsg_server_replay_progress(tsg_server *s, int n, const int *handles, const int *ops,
                                          const double *arrival_ms, int keep, unsigned long long *ids,
                                          tsg_request_info *infos, tsg_replay_progress progress, void *user) {
    const double t0 = clock_ms();
    int first_err = TSG_OK;
    for (int i = 0; i < n; i++) {
        const double due = t0 + arrival_ms[i];
        double now;
        while ((now = clock_ms()) < due) {
            const double d = due - now;
            if (d > 0.2) std::this_thread::sleep_for(std::chrono::microseconds((long long)((d - 0.1) * 1000)));
        }
        int done = 0;
        {
            std::lock_guard<std::mutex> g(s->mu);
            const int st = submit_locked(s, handles[i], ops[i], !keep, &ids[i]);
            if (st) { ids[i] = 0; if (!first_err) first_err = st; }
            if (progress)
                for (int j = 0; j <= i; j++) {
                    auto it = ids[j] ? s->reqs.find(ids[j]) : s->reqs.end();
                    done += it != s->reqs.end() && it->second->done;
                }
        }
        if (progress) progress(i + 1, done, user);
    }
    for (int i = 0; i < n; i++) {
        if (ids[i]) {
            const int st = tsg_server_wait(s, ids[i], 0, infos ? &infos[i] : nullptr);
            if (st && !first_err) first_err = st;
            if (!keep) {
                std::lock_guard<std::mutex> g(s->mu);
                s->reqs.erase(ids[i]);
            }
        }
        if (progress) progress(n, i + 1, user);
    }
    return first_err;
}

extern "C" void tsg_server_destroy(tsg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        // pending requests are dropped; results go back to their workers
        s->queue.clear();
        for (auto &kv : s->reqs) {
            Request *r = kv.second.get();
            if (r->done && !r->released) free_request_results(s, r);
        }
        s->cv_work.notify_all();
    }
    for (auto &t : s->workers) t.join();
    for (int w = 0; w < (int)s->wctx.size(); w++) {
        for (auto &t : s->pending_free[w]) tsg_tiles_free(s->wctx[w], &t);
        tsg_ctx_destroy(s->wctx[w]);
    }
    for (auto &M : s->mats) {
        tsg_tiles_free(s->main, &M->tat);
        tsg_tiles_free(s->main, &M->ta);
        tsg_dcsr_free(s->main, &M->at);
        tsg_dcsr_free(s->main, &M->a);
    }
    tsg_ctx_destroy(s->main);
    tsg_arena_destroy(s->arena);
    delete s;
}
