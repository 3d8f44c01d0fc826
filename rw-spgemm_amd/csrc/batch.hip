// batch.hip -- many independent SpGEMMs per step in fused launches.
//
// A batch of problems C_i = A_i * B_i (B_i = A_i or A_i^T) is packed into
// one block-diagonal product
//
//     blockdiag(A_0, A_1, ...) * blockdiag(B_0, B_1, ...) = blockdiag(C_i)
//
// with every block starting on a 16-row / 16-column boundary, so the tile
// structure of the packed operands is exactly the concatenation of the
// problems' own tile structures (no tile straddles two problems) and each
// pipeline kernel runs once for the whole batch: one block per tile-row of
// any problem, the small problems filling the GPU together.  Every C tile
// of problem i is computed from problem i's tiles only, in the same
// ascending-k order as a single call, so each C_i is bit-identical to
// tsg_spgemm on (A_i, B_i).
//
// The first (planning) run is the synchronous pipeline; it fixes C's sizes
// and the kernel choices.  The enqueue-only form of the same pipeline is
// then captured once into a HIP graph, and a step is one graph launch that
// recomputes steps 1-3 of every problem into the resident C buffers.
// Packed products whose C would overflow the int32 offsets of the tile
// format are split into consecutive groups, each a fused pipeline of its
// own inside the same graph.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "tsg_internal.h"

namespace {

// One block of a packed CSR: rows [0, m) of a source CSR (rowptr points
// at the block's first row; its values index colidx/val absolutely),
// padded to mpad rows and placed at row dst_row / nonzero dst_nnz of the
// packed matrix with columns shifted by colshift.
struct PackDesc {
    const int *rowptr;
    const int *colidx;
    const double *val;
    int m, mpad, nnz;
    int dst_row, dst_nnz, colshift;
};

__device__ __forceinline__ int find_block(const int *__restrict__ base, int count, int x) {
    // last block b with base[b] <= x (base ascending, base[0] = 0)
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256)
k_pack_rowptr(const PackDesc *__restrict__ d, const int *__restrict__ rowbase, int count, int M, int total,
              int *__restrict__ out) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r > M) return;
    if (r == M) { out[M] = total; return; }
    const int b = find_block(rowbase, count, r);
    const PackDesc &p = d[b];
    const int lr = r - p.dst_row;
    out[r] = p.dst_nnz + (lr < p.m ? p.rowptr[lr] - p.rowptr[0] : p.nnz);
}

__global__ void __launch_bounds__(256)
k_pack_entries(const PackDesc *__restrict__ d, const int *__restrict__ nnzbase, int count, long long total,
               int *__restrict__ col, double *__restrict__ val) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int b = find_block(nnzbase, count, (int)e);
    const PackDesc &p = d[b];
    const int s = p.rowptr[0] + (int)(e - p.dst_nnz);
    col[e] = p.colidx[s] + p.colshift;
    val[e] = p.val[s];
}

int pad16(int x) { return (x + 15) & ~15; }

// Packs the blocks (host descriptors; dst_row/dst_nnz filled here) into a
// new device CSR of M x N.
int pack_csr(tsg_ctx *ctx, std::vector<PackDesc> &ds, int N, tsg_dcsr *out) {
    memset(out, 0, sizeof(*out));
    const int count = (int)ds.size();
    long long M = 0, nnz = 0;
    std::vector<int> rowbase(count), nnzbase(count);
    for (int i = 0; i < count; i++) {
        rowbase[i] = (int)M;
        nnzbase[i] = (int)nnz;
        ds[i].dst_row = (int)M;
        ds[i].dst_nnz = (int)nnz;
        M += ds[i].mpad;
        nnz += ds[i].nnz;
        if (M > 0x7fffffffLL - 16 || nnz > 0x7fffffffLL) {
            tsg_set_error("batch: packed operand exceeds int32 rows or nonzeros");
            return TSG_ERR_OVERFLOW;
        }
    }
    out->m = (int)M;
    out->n = N;
    out->nnz = (int)nnz;
    PackDesc *dd = nullptr;
    int *drb = nullptr, *dnb = nullptr;
    TSG_TRY(tsg_alloc_t(ctx, &dd, (size_t)count));
    TSG_TRY(tsg_alloc_t(ctx, &drb, (size_t)count));
    TSG_TRY(tsg_alloc_t(ctx, &dnb, (size_t)count));
    TSG_HIP(hipMemcpyAsync(dd, ds.data(), sizeof(PackDesc) * count, hipMemcpyHostToDevice, ctx->stream));
    TSG_HIP(hipMemcpyAsync(drb, rowbase.data(), 4 * (size_t)count, hipMemcpyHostToDevice, ctx->stream));
    TSG_HIP(hipMemcpyAsync(dnb, nnzbase.data(), 4 * (size_t)count, hipMemcpyHostToDevice, ctx->stream));
    TSG_TRY(tsg_alloc_t(ctx, &out->rowptr, (size_t)M + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->colidx, (size_t)nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->val, (size_t)nnz + 1));
    hipLaunchKernelGGL(k_pack_rowptr, dim3(tsg_div_up(M + 1, 256)), dim3(256), 0, ctx->stream,
                       (const PackDesc *)dd, (const int *)drb, count, (int)M, (int)nnz, out->rowptr);
    TSG_LAUNCH_CHECK();
    if (nnz > 0) {
        hipLaunchKernelGGL(k_pack_entries, dim3(tsg_div_up(nnz, 256)), dim3(256), 0, ctx->stream,
                           (const PackDesc *)dd, (const int *)dnb, count, nnz, out->colidx, out->val);
        TSG_LAUNCH_CHECK();
    }
    TSG_HIP(hipStreamSynchronize(ctx->stream));  // host vectors go out of scope
    tsg_free(ctx, dd);
    tsg_free(ctx, drb);
    tsg_free(ctx, dnb);
    return TSG_OK;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// One fused pipeline over a consecutive range of problems.
struct tsg_batch_group {
    int first, count;
    tsg_dtiles A, B, C;
    bool b_is_a;
    tsg_spgemm_plan plan;
    void *work;
};

struct tsg_batch {
    tsg_ctx *ctx;
    int count;
    std::vector<int> m, n, op, nnz;
    std::vector<int> group_of, rowbase, cbase;  // per problem, inside its group
    std::vector<tsg_batch_group> groups;
    tsg_tune tune;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int graph_nodes;
    double build_ms, plan_ms, capture_ms;
};

static void free_group(tsg_ctx *ctx, tsg_batch_group &g) {
    tsg_tiles_free(ctx, &g.C);
    if (!g.b_is_a) tsg_tiles_free(ctx, &g.B);
    tsg_tiles_free(ctx, &g.A);
    tsg_free(ctx, g.work);
    g.work = nullptr;
}

static void drop_graph(tsg_batch *b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Packs problems [first, first+cnt) and converts them to tiles.
static int build_group(tsg_batch *b, const tsg_dcsr *probs, int first, int cnt, tsg_batch_group &g) {
    tsg_ctx *ctx = b->ctx;
    memset(&g.A, 0, sizeof(g.A));
    memset(&g.B, 0, sizeof(g.B));
    memset(&g.C, 0, sizeof(g.C));
    g.first = first;
    g.count = cnt;
    g.work = nullptr;
    bool all_a2 = true;
    std::vector<PackDesc> da(cnt);
    int Ncols = 0;
    for (int k = 0; k < cnt; k++) {
        const int i = first + k;
        const tsg_dcsr &p = probs[i];
        da[k] = PackDesc{p.rowptr, p.colidx, p.val, p.m, pad16(p.m), p.nnz, 0, 0, Ncols};
        Ncols += pad16(p.n);
        if (b->op[i] != TSG_OP_A2) all_a2 = false;
    }
    tsg_dcsr pa, pat, pb;
    TSG_TRY(pack_csr(ctx, da, Ncols, &pa));
    // tile-aligned bases of each problem: A rows, A columns (= B rows), C columns
    std::vector<int> arow(cnt), acol(cnt), ccol(cnt);
    int cb = 0;
    for (int k = 0; k < cnt; k++) {
        const int i = first + k;
        arow[k] = da[k].dst_row;
        acol[k] = da[k].colshift;
        ccol[k] = cb;
        cb += pad16(b->op[i] == TSG_OP_A2 ? b->n[i] : b->m[i]);
        b->group_of[i] = (int)b->groups.size();
        b->rowbase[i] = arow[k];
        b->cbase[i] = ccol[k];
    }
    int st = TSG_OK;
    g.b_is_a = all_a2;
    if (!all_a2) {
        memset(&pat, 0, sizeof(pat));
        st = tsg_dcsr_transpose(ctx, &pa, &pat);
        if (st) { tsg_dcsr_free(ctx, &pa); return st; }
        // B block k: rows acol[k].., from A (A^2) or A^T (AA^T), columns moved to ccol[k]
        std::vector<PackDesc> dbv(cnt);
        for (int k = 0; k < cnt; k++) {
            const int i = first + k;
            if (b->op[i] == TSG_OP_A2)
                dbv[k] = PackDesc{pa.rowptr + arow[k], pa.colidx, pa.val, b->m[i], pad16(b->n[i]), b->nnz[i],
                                  0, 0, ccol[k] - acol[k]};
            else  // rows of A_i^T sit at acol[k] in the transposed packed matrix
                dbv[k] = PackDesc{pat.rowptr + acol[k], pat.colidx, pat.val, b->n[i], pad16(b->n[i]), b->nnz[i],
                                  0, 0, ccol[k] - arow[k]};
        }
        if (!st) st = pack_csr(ctx, dbv, cb, &pb);
        tsg_dcsr_free(ctx, &pat);
        if (st) { tsg_dcsr_free(ctx, &pa); return st; }
    }
    st = tsg_csr2tile(ctx, &pa, &g.A);
    if (!st && all_a2) {
        st = tsg_tiles_build_csc(ctx, &g.A);
        if (!st) st = tsg_tiles_build_rowindex(ctx, &g.A);
        g.B = g.A;
    } else if (!st) {
        st = tsg_tiles_build_rowindex(ctx, &g.A);  // the element kernels walk A's rows
        if (!st) st = tsg_csr2tile(ctx, &pb, &g.B);
        if (!st) st = tsg_tiles_build_csc(ctx, &g.B);
        if (!st) st = tsg_tiles_build_rowindex(ctx, &g.B);
        tsg_dcsr_free(ctx, &pb);
    }
    tsg_dcsr_free(ctx, &pa);
    return st;
}

// Planning run + work buffer for one group under the batch's tune.
static int plan_group(tsg_batch *b, tsg_batch_group &g) {
    tsg_ctx *ctx = b->ctx;
    tsg_tiles_free(ctx, &g.C);
    tsg_free(ctx, g.work);
    g.work = nullptr;
    TSG_TRY(tsg_spgemm_planned(ctx, &g.A, &g.B, &b->tune, &g.C, nullptr, &g.plan));
    if (b->tune.batch_group_nnz_cap > 0 && g.plan.nnzC > b->tune.batch_group_nnz_cap) {
        tsg_set_error("batch: group C nonzeros %lld above the cap %d", g.plan.nnzC, b->tune.batch_group_nnz_cap);
        return TSG_ERR_OVERFLOW;
    }
    TSG_TRY(tsg_alloc(ctx, &g.work, tsg_spgemm_work_bytes(&g.C, &g.B, g.plan.s1w)));
    return TSG_OK;
}

static int capture(tsg_batch *b) {
    drop_graph(b);
    tsg_ctx *ctx = b->ctx;
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    const double t0 = now_ms();
    TSG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int st = TSG_OK;
    for (auto &g : b->groups) {
        st = tsg_spgemm_enqueue(ctx->stream, &g.A, &g.B, &g.C, g.plan, g.work);
        if (st) break;
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st) { if (graph) hipGraphDestroy(graph); return st; }
    if (e != hipSuccess) { tsg_set_error("batch: graph capture failed: %s", hipGetErrorString(e)); return TSG_ERR_HIP; }
    b->graph = graph;
    TSG_HIP(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
    size_t nn = 0;
    TSG_HIP(hipGraphGetNodes(graph, nullptr, &nn));
    b->graph_nodes = (int)nn;
    b->capture_ms = now_ms() - t0;
    return TSG_OK;
}

extern "C" void tsg_batch_destroy(tsg_batch *b) {
    if (!b) return;
    hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (auto &g : b->groups) free_group(b->ctx, g);
    delete b;
}

// Builds groups over [first, first+cnt), halving a range whose packed
// product overflows the int32 offsets.
static int build_range(tsg_batch *b, const tsg_dcsr *probs, int first, int cnt) {
    tsg_batch_group g;
    int st = build_group(b, probs, first, cnt, g);
    if (st == TSG_OK) {
        b->groups.push_back(g);
        st = plan_group(b, b->groups.back());
        if (st == TSG_OK) return TSG_OK;
        tsg_batch_group bad = b->groups.back();
        b->groups.pop_back();
        free_group(b->ctx, bad);
    }
    if (st != TSG_ERR_OVERFLOW || cnt == 1) return st;
    const int h = cnt / 2;
    TSG_TRY(build_range(b, probs, first, h));
    return build_range(b, probs, first + h, cnt - h);
}

extern "C" int tsg_batch_create(tsg_ctx *ctx, const tsg_dcsr *probs, const int *ops, int count, const tsg_tune *tune,
                                tsg_batch **out) {
    *out = nullptr;
    if (!ctx || !probs || !ops || count <= 0) { tsg_set_error("tsg_batch_create: empty batch"); return TSG_ERR_ARG; }
    for (int i = 0; i < count; i++) {
        const tsg_dcsr &p = probs[i];
        if (ops[i] != TSG_OP_A2 && ops[i] != TSG_OP_AAT) {
            tsg_set_error("tsg_batch_create: problem %d: op must be TSG_OP_A2 or TSG_OP_AAT", i);
            return TSG_ERR_ARG;
        }
        if (ops[i] == TSG_OP_A2 && p.m != p.n) {
            tsg_set_error("tsg_batch_create: problem %d: A^2 needs a square matrix (%d x %d)", i, p.m, p.n);
            return TSG_ERR_ARG;
        }
        if (p.m < 0 || p.n < 0 || p.nnz < 0 || !p.rowptr) {
            tsg_set_error("tsg_batch_create: problem %d: bad CSR", i);
            return TSG_ERR_ARG;
        }
        TSG_TRY(tsg_validate_csr(ctx, &p));
    }
    tsg_batch *b = new tsg_batch();
    b->ctx = ctx;
    b->count = count;
    b->m.resize(count); b->n.resize(count); b->op.resize(count); b->nnz.resize(count);
    b->group_of.resize(count); b->rowbase.resize(count); b->cbase.resize(count);
    for (int i = 0; i < count; i++) {
        b->m[i] = probs[i].m; b->n[i] = probs[i].n; b->op[i] = ops[i]; b->nnz[i] = probs[i].nnz;
    }
    if (tune) b->tune = *tune; else tsg_tune_default(&b->tune);
    b->graph = nullptr;
    b->exec = nullptr;
    const double t0 = now_ms();
    int st = build_range(b, probs, 0, count);
    if (st == TSG_OK) {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) { tsg_set_error("batch: %s", hipGetErrorString(e)); st = TSG_ERR_HIP; }
    }
    b->build_ms = now_ms() - t0;
    b->plan_ms = b->build_ms;
    if (st == TSG_OK) st = capture(b);
    if (st) { tsg_batch_destroy(b); return st; }
    *out = b;
    return TSG_OK;
}

extern "C" int tsg_batch_run(tsg_batch *b) {
    if (!b || !b->exec) { tsg_set_error("tsg_batch_run: batch not ready"); return TSG_ERR_STATE; }
    TSG_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    return TSG_OK;
}

// Switches the batch to `t`: re-plans every group and re-captures.
static int retune(tsg_batch *b, const tsg_tune &t) {
    b->tune = t;
    drop_graph(b);
    for (auto &g : b->groups) TSG_TRY(plan_group(b, g));
    return capture(b);
}

static int time_graph(tsg_batch *b, int reps, double *best_ms) {
    hipEvent_t e0, e1;
    TSG_HIP(hipEventCreate(&e0));
    TSG_HIP(hipEventCreate(&e1));
    double best = 1e30;
    for (int r = 0; r < reps; r++) {
        TSG_HIP(hipEventRecord(e0, b->ctx->stream));
        TSG_TRY(tsg_batch_run(b));
        TSG_HIP(hipEventRecord(e1, b->ctx->stream));
        TSG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        TSG_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, (double)ms);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *best_ms = best;
    return TSG_OK;
}

// Autotuner: times each candidate configuration over whole graph launches
// of this batch and keeps the fastest.  Candidates: the step-2/3 strategy
// (row-wise Gustavson enumeration in LDS vs tile-parallel list
// intersection) and the dense-accumulator threshold of the tile-parallel
// numeric kernel.  Results do not depend on the choice (every kernel sums
// in the same order), only the time does.
extern "C" int tsg_batch_autotune(tsg_batch *b, int reps, tsg_tune *chosen, double *best_ms) {
    if (!b) { tsg_set_error("tsg_batch_autotune: null batch"); return TSG_ERR_ARG; }
    if (reps < 1) reps = 1;
    std::vector<tsg_tune> cands;
    auto add = [&](int fp, int em, int th) {
        tsg_tune t;
        tsg_tune_default(&t);
        t.force_tile_parallel = fp;
        t.element_mode = em;
        t.dense_acc_min_nnz = th;
        t.batch_group_nnz_cap = b->tune.batch_group_nnz_cap;
        cands.push_back(t);
    };
    // element-wise numeric step for the rows the row-wise kernel cannot
    // hold, or for every row; the tile-parallel kernel with several
    // dense-accumulator thresholds for those rows, or for every row
    long long tiles = 0;
    for (const auto &g : b->groups) tiles += g.plan.numtileC;
    add(0, 0, 64);
    add(0, 2, 64);
    for (int th : {16, 64, 257}) add(0, 1, th);
    // every tile-row tile-parallel only while that stays affordable to try
    if (tiles < (64LL << 20))
        for (int th : {64, 257}) add(1, 1, th);
    double best = 1e30;
    tsg_tune bt = b->tune;
    for (const auto &t : cands) {
        TSG_TRY(retune(b, t));
        double ms;
        TSG_TRY(time_graph(b, 1, &ms));  // first launch; far slower candidates stop here
        if (ms < 4.0 * best) TSG_TRY(time_graph(b, reps, &ms));
        if (ms < best) { best = ms; bt = t; }
    }
    TSG_TRY(retune(b, bt));
    if (chosen) *chosen = bt;
    if (best_ms) *best_ms = best;
    return TSG_OK;
}

extern "C" int tsg_batch_get_info(const tsg_batch *b, tsg_batch_info *out) {
    memset(out, 0, sizeof(*out));
    if (!b) return TSG_ERR_ARG;
    out->count = b->count;
    out->groups = (int)b->groups.size();
    for (const auto &g : b->groups) {
        out->numtileC += g.plan.numtileC;
        out->nnzC += g.plan.nnzC;
    }
    out->graph_nodes = b->graph_nodes;
    out->build_ms = b->build_ms;
    out->capture_ms = b->capture_ms;
    out->tune = b->tune;
    return TSG_OK;
}

// Problem i's C as a tile-row view of its group's packed C.
static int problem_view(tsg_batch *b, int i, tsg_dtiles *v, int *jbase) {
    const tsg_batch_group &g = b->groups[b->group_of[i]];
    const int r0 = b->rowbase[i] / 16;
    *v = g.C;
    v->tile_ptr = g.C.tile_ptr + r0;
    v->m = b->m[i];
    v->tilem = (b->m[i] + 15) / 16;
    const int nc = b->op[i] == TSG_OP_A2 ? b->n[i] : b->m[i];
    v->n = nc;
    v->tilen = (nc + 15) / 16;
    *jbase = b->cbase[i] / 16;
    return TSG_OK;
}

template <typename T>
static int dl_range(tsg_ctx *ctx, T **dst, const T *src, long long n) {
    *dst = (T *)malloc((size_t)(n > 0 ? n : 1) * sizeof(T));
    if (!*dst) { tsg_set_error("host allocation failed"); return TSG_ERR_NOMEM; }
    if (n > 0) TSG_HIP(hipMemcpyAsync(*dst, src, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    return TSG_OK;
}

extern "C" int tsg_batch_result(tsg_batch *b, int i, tsg_htiles *tiles, tsg_hcsr *csr) {
    if (!b || i < 0 || i >= b->count) { tsg_set_error("tsg_batch_result: bad problem index"); return TSG_ERR_ARG; }
    tsg_ctx *ctx = b->ctx;
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    tsg_dtiles v;
    int jbase;
    TSG_TRY(problem_view(b, i, &v, &jbase));
    if (tiles) {
        memset(tiles, 0, sizeof(*tiles));
        int tp[2], nz[2];
        TSG_TRY(tsg_d2h(ctx, &tp[0], v.tile_ptr, 4));
        TSG_TRY(tsg_d2h(ctx, &tp[1], v.tile_ptr + v.tilem, 4));
        TSG_TRY(tsg_d2h(ctx, &nz[0], v.tile_nnz + tp[0], 4));
        TSG_TRY(tsg_d2h(ctx, &nz[1], v.tile_nnz + tp[1], 4));
        const long long nt = tp[1] - tp[0], nn = nz[1] - nz[0];
        tiles->m = v.m; tiles->n = v.n; tiles->tilem = v.tilem; tiles->tilen = v.tilen;
        tiles->numtile = (int)nt;
        tiles->nnz = (int)nn;
        TSG_TRY(dl_range(ctx, &tiles->tile_ptr, v.tile_ptr, v.tilem + 1));
        TSG_TRY(dl_range(ctx, &tiles->tile_colidx, v.tile_colidx + tp[0], nt));
        TSG_TRY(dl_range(ctx, &tiles->tile_rowidx, v.tile_rowidx + tp[0], nt));
        TSG_TRY(dl_range(ctx, &tiles->tile_nnz, v.tile_nnz + tp[0], nt + 1));
        TSG_TRY(dl_range(ctx, &tiles->csr_ptr, v.csr_ptr + 16 * (long long)tp[0], 16 * nt));
        TSG_TRY(dl_range(ctx, &tiles->mask, v.mask + 16 * (long long)tp[0], 16 * nt));
        TSG_TRY(dl_range(ctx, &tiles->csr_col, v.csr_col + nz[0], nn));
        TSG_TRY(dl_range(ctx, &tiles->val, v.val + nz[0], nn));
        TSG_HIP(hipStreamSynchronize(ctx->stream));
        const int r0 = b->rowbase[i] / 16;
        for (int r = 0; r <= v.tilem; r++) tiles->tile_ptr[r] -= tp[0];
        for (long long t = 0; t < nt; t++) {
            tiles->tile_colidx[t] -= jbase;
            tiles->tile_rowidx[t] -= r0;
        }
        for (long long t = 0; t <= nt; t++) tiles->tile_nnz[t] -= nz[0];
    }
    if (csr) {
        tsg_dcsr d;
        TSG_TRY(tsg_tile2csr_view(ctx, &v, jbase, &d));
        int st = tsg_dcsr_download(ctx, &d, csr);
        tsg_dcsr_free(ctx, &d);
        TSG_TRY(st);
    }
    return TSG_OK;
}

extern "C" int tsg_batch_result_checksum(tsg_batch *b, int i, tsg_csum *out) {
    if (!b || i < 0 || i >= b->count) {
        tsg_set_error("tsg_batch_result_checksum: bad problem index");
        return TSG_ERR_ARG;
    }
    TSG_HIP(hipStreamSynchronize(b->ctx->stream));
    tsg_dtiles v;
    int jbase;
    TSG_TRY(problem_view(b, i, &v, &jbase));
    return tsg_tiles_checksum(b->ctx, &v, 0, jbase, out);
}

// ------------------------------------------------------------------ tile packing
// Block-diagonal packing of resident tile matrices (no CSR round trip):
// part p lands at tile-row TR_p, tile-column TC_p, tile TB_p and nonzero
// NB_p of the packed matrix; every index array is shifted, every data
// array copied.  Used by the server to fuse queued requests whose operands
// are already in HBM as tiles.
namespace {
struct TilePart {
    const int *tile_ptr, *tile_colidx, *tile_nnz, *csc_ptr, *csc_rowidx, *csc_tileid, *rowx_ptr, *rowx_col;
    const unsigned char *csr_ptr, *csr_col;
    const unsigned short *mask;
    const double *val, *rowx_val;
    int m, tilem, tilen, numtile, nnz;
    int TR, TC, TB, NB;
};

__global__ void __launch_bounds__(256)
k_tp_rows(const TilePart *__restrict__ P, const int *__restrict__ trb, int count, int tilem, int total_tiles,
          int *__restrict__ tile_ptr) {
    const int I = blockIdx.x * 256 + threadIdx.x;
    if (I > tilem) return;
    if (I == tilem) { tile_ptr[tilem] = total_tiles; return; }
    const TilePart &p = P[find_block(trb, count, I)];
    tile_ptr[I] = p.TB + p.tile_ptr[I - p.TR];
}

__global__ void __launch_bounds__(256)
k_tp_cols(const TilePart *__restrict__ P, const int *__restrict__ tcb, int count, int tilen, int total_tiles,
          int *__restrict__ csc_ptr) {
    const int J = blockIdx.x * 256 + threadIdx.x;
    if (J > tilen) return;
    if (J == tilen) { csc_ptr[tilen] = total_tiles; return; }
    const TilePart &p = P[find_block(tcb, count, J)];
    const int lj = J - p.TC;
    csc_ptr[J] = p.TB + (lj < p.tilen ? p.csc_ptr[lj] : p.numtile);
}

__global__ void __launch_bounds__(256)
k_tp_tiles(const TilePart *__restrict__ P, const int *__restrict__ tbb, int count, int total_tiles,
           int *__restrict__ colidx, int *__restrict__ rowidx, int *__restrict__ tnnz, int *__restrict__ cscrow,
           int *__restrict__ cscid, uint4 *__restrict__ ptr16, uint4 *__restrict__ mask16) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total_tiles) return;
    const TilePart &p = P[find_block(tbb, count, t)];
    const int lt = t - p.TB;
    colidx[t] = p.tile_colidx[lt] + p.TC;
    tnnz[t] = p.tile_nnz[lt] + p.NB;
    if (rowidx) {
        // tile row of local tile lt: last local tile-row whose start <= lt
        int lo = 0, hi = p.tilem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.tile_ptr[mid] <= lt) lo = mid; else hi = mid - 1;
        }
        rowidx[t] = lo + p.TR;
    }
    if (cscrow) {
        cscrow[t] = p.csc_rowidx[lt] + p.TR;
        cscid[t] = p.csc_tileid[lt] + p.TB;
    }
    ptr16[t] = reinterpret_cast<const uint4 *>(p.csr_ptr)[lt];
    mask16[2 * t] = reinterpret_cast<const uint4 *>(p.mask)[2 * lt];
    mask16[2 * t + 1] = reinterpret_cast<const uint4 *>(p.mask)[2 * lt + 1];
}

__global__ void __launch_bounds__(256)
k_tp_nnz(const TilePart *__restrict__ P, const int *__restrict__ nbb, int count, long long total_nnz,
         unsigned char *__restrict__ col8, double *__restrict__ val, int *__restrict__ xcol, double *__restrict__ xval) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total_nnz) return;
    const TilePart &p = P[find_block(nbb, count, (int)e)];
    const int le = (int)(e - p.NB);
    col8[e] = p.csr_col[le];
    val[e] = p.val[le];
    if (xcol) {
        xcol[e] = p.rowx_col[le] + 16 * p.TC;
        xval[e] = p.rowx_val[le];
    }
}

__global__ void __launch_bounds__(256)
k_tp_xrows(const TilePart *__restrict__ P, const int *__restrict__ trb, int count, int m, int total_nnz,
           int *__restrict__ xptr) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i > m) return;
    if (i == m) { xptr[m] = total_nnz; return; }
    const TilePart &p = P[find_block(trb, count, i >> 4)];
    const int li = i - 16 * p.TR;
    xptr[i] = p.NB + (li < p.m ? p.rowx_ptr[li] : p.nnz);
}
}  // namespace

// parts[i] packed at tile-row/tile-column bases; out gets the CSC index
// and the row index when every part has them.  The packed matrix has
// 16 * sum(tilem) rows and 16 * sum(tilen) columns.
int tsg_tiles_pack(tsg_ctx *ctx, const tsg_dtiles *const *parts, int count, tsg_dtiles *out) {
    memset(out, 0, sizeof(*out));
    if (count <= 0) { tsg_set_error("tsg_tiles_pack: no parts"); return TSG_ERR_ARG; }
    std::vector<TilePart> P(count);
    std::vector<int> trb(count), tcb(count), tbb(count), nbb(count);
    long long TR = 0, TC = 0, TB = 0, NB = 0;
    bool csc = true, rowx = true;
    for (int i = 0; i < count; i++) {
        const tsg_dtiles *t = parts[i];
        csc = csc && t->csc_ptr;
        rowx = rowx && t->rowx_ptr;
        P[i] = TilePart{t->tile_ptr, t->tile_colidx, t->tile_nnz, t->csc_ptr, t->csc_rowidx, t->csc_tileid,
                        t->rowx_ptr, t->rowx_col, t->csr_ptr, t->csr_col, t->mask, t->val, t->rowx_val,
                        t->m, t->tilem, t->tilen, t->numtile, t->nnz,
                        (int)TR, (int)TC, (int)TB, (int)NB};
        trb[i] = (int)TR; tcb[i] = (int)TC; tbb[i] = (int)TB; nbb[i] = (int)NB;
        TR += t->tilem; TC += t->tilen; TB += t->numtile; NB += t->nnz;
        if (16 * TR > 0x7fffffffLL || 16 * TC > 0x7fffffffLL || TB > 0x7fffffffLL || NB > 0x7fffffffLL) {
            tsg_set_error("tsg_tiles_pack: packed matrix exceeds int32 indices");
            return TSG_ERR_OVERFLOW;
        }
    }
    out->tilem = (int)TR; out->tilen = (int)TC; out->m = (int)(16 * TR); out->n = (int)(16 * TC);
    out->numtile = (int)TB; out->nnz = (int)NB;
    TilePart *dP;
    int *dtrb, *dtcb, *dtbb, *dnbb;
    hipStream_t s = ctx->stream;
    TSG_TRY(tsg_alloc_t(ctx, &dP, (size_t)count));
    TSG_TRY(tsg_alloc_t(ctx, &dtrb, (size_t)count));
    TSG_TRY(tsg_alloc_t(ctx, &dtcb, (size_t)count));
    TSG_TRY(tsg_alloc_t(ctx, &dtbb, (size_t)count));
    TSG_TRY(tsg_alloc_t(ctx, &dnbb, (size_t)count));
    TSG_HIP(hipMemcpyAsync(dP, P.data(), sizeof(TilePart) * count, hipMemcpyHostToDevice, s));
    TSG_HIP(hipMemcpyAsync(dtrb, trb.data(), 4 * (size_t)count, hipMemcpyHostToDevice, s));
    TSG_HIP(hipMemcpyAsync(dtcb, tcb.data(), 4 * (size_t)count, hipMemcpyHostToDevice, s));
    TSG_HIP(hipMemcpyAsync(dtbb, tbb.data(), 4 * (size_t)count, hipMemcpyHostToDevice, s));
    TSG_HIP(hipMemcpyAsync(dnbb, nbb.data(), 4 * (size_t)count, hipMemcpyHostToDevice, s));
    TSG_TRY(tsg_alloc_t(ctx, &out->tile_ptr, (size_t)TR + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->tile_colidx, (size_t)TB + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->tile_rowidx, (size_t)TB + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->tile_nnz, (size_t)TB + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->csr_ptr, (size_t)TB * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &out->mask, (size_t)TB * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &out->csr_col, (size_t)NB + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->val, (size_t)NB + 1));
    if (csc) {
        TSG_TRY(tsg_alloc_t(ctx, &out->csc_ptr, (size_t)TC + 1));
        TSG_TRY(tsg_alloc_t(ctx, &out->csc_rowidx, (size_t)TB + 1));
        TSG_TRY(tsg_alloc_t(ctx, &out->csc_tileid, (size_t)TB + 1));
    }
    if (rowx) {
        TSG_TRY(tsg_alloc_t(ctx, &out->rowx_ptr, (size_t)out->m + 1));
        TSG_TRY(tsg_alloc_t(ctx, &out->rowx_col, (size_t)NB + 1));
        TSG_TRY(tsg_alloc_t(ctx, &out->rowx_val, (size_t)NB + 1));
    }
    hipLaunchKernelGGL(k_tp_rows, dim3(tsg_div_up(TR + 1, 256)), dim3(256), 0, s, (const TilePart *)dP,
                       (const int *)dtrb, count, (int)TR, (int)TB, out->tile_ptr);
    TSG_LAUNCH_CHECK();
    // the tile-row of every tile (k_tp_tiles searches each part's tile_ptr)
    if (TB > 0) {
        hipLaunchKernelGGL(k_tp_tiles, dim3(tsg_div_up(TB, 256)), dim3(256), 0, s, (const TilePart *)dP,
                           (const int *)dtbb, count, (int)TB, out->tile_colidx, out->tile_rowidx, out->tile_nnz,
                           csc ? out->csc_rowidx : (int *)nullptr, csc ? out->csc_tileid : (int *)nullptr,
                           reinterpret_cast<uint4 *>(out->csr_ptr), reinterpret_cast<uint4 *>(out->mask));
        TSG_LAUNCH_CHECK();
    }
    // tile_nnz[TB] = NB
    TSG_HIP(hipMemcpyAsync(out->tile_nnz + TB, &out->nnz, 4, hipMemcpyHostToDevice, s));
    if (csc) {
        hipLaunchKernelGGL(k_tp_cols, dim3(tsg_div_up(TC + 1, 256)), dim3(256), 0, s, (const TilePart *)dP,
                           (const int *)dtcb, count, (int)TC, (int)TB, out->csc_ptr);
        TSG_LAUNCH_CHECK();
    }
    if (NB > 0) {
        hipLaunchKernelGGL(k_tp_nnz, dim3(tsg_div_up(NB, 256)), dim3(256), 0, s, (const TilePart *)dP,
                           (const int *)dnbb, count, NB, out->csr_col, out->val,
                           rowx ? out->rowx_col : (int *)nullptr, rowx ? out->rowx_val : (double *)nullptr);
        TSG_LAUNCH_CHECK();
    }
    if (rowx) {
        hipLaunchKernelGGL(k_tp_xrows, dim3(tsg_div_up((long long)out->m + 1, 256)), dim3(256), 0, s,
                           (const TilePart *)dP, (const int *)dtrb, count, out->m, (int)NB, out->rowx_ptr);
        TSG_LAUNCH_CHECK();
    }
    TSG_HIP(hipStreamSynchronize(s));  // host descriptors and out->nnz go out of scope
    tsg_free(ctx, dP);
    tsg_free(ctx, dtrb);
    tsg_free(ctx, dtcb);
    tsg_free(ctx, dtbb);
    tsg_free(ctx, dnbb);
    return TSG_OK;
}
