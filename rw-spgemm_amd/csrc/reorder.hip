// reorder.hip -- matrix orderings and permutations on the device.
//
//  * degree order: vertices sorted by (-degree, id)
//    (reorder_algorithm/common/order_utils.py:78-81)
//  * Cuthill-McKee / RCM on the symmetrised pattern, the structure of the
//    reference's native_rcm (order_utils.py:84-104) with every tie broken by
//    (degree, id): components are labelled by their lowest-(degree, id)
//    vertex, all components are traversed by one multi-source
//    level-synchronous BFS (a vertex is claimed by its first parent in
//    level order, children ordered by (parent, degree, id)), and the
//    concatenated levels are grouped by component and reversed.
//  * seeded random order
//  * P*A*Q^T with sorted rows: apply_bipermutation_csr / restore
//    (src/reorder.h:2429-2551), done with the library's radix sort.
#include <string.h>

#include "tsg_internal.h"

static int bits_of(long long v) {
    int b = 1;
    while ((1ll << b) <= v) b++;
    return b;
}

__global__ void k_iota2(int *v, long long n) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (int)i;
}
__global__ void k_scatter_inverse(const int *__restrict__ old_id, int n, int *__restrict__ new_id) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) new_id[old_id[i]] = i;
}
__global__ void k_gather_i32_r(const int *__restrict__ src, const int *__restrict__ idx, int n, int *__restrict__ dst) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_perm_check(const int *__restrict__ p, int n, int *__restrict__ seen, int *__restrict__ bad) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int v = p[i];
    if (v < 0 || v >= n) { atomicOr(bad, 1); return; }
    if (atomicAdd(&seen[v], 1) != 0) atomicOr(bad, 1);
}

// TSG_OK when d_perm[0..n) is a permutation of 0..n-1.
extern "C" int tsg_perm_validate(tsg_ctx *ctx, int n, const int *d_perm) {
    if (n <= 0) return TSG_OK;
    int *seen;
    TSG_TRY(tsg_alloc_t(ctx, &seen, (size_t)n + 1));
    TSG_HIP(hipMemsetAsync(seen, 0, ((size_t)n + 1) * 4, ctx->stream));
    hipLaunchKernelGGL(k_perm_check, dim3(tsg_div_up(n, 256)), dim3(256), 0, ctx->stream, d_perm, n, seen, seen + n);
    TSG_LAUNCH_CHECK();
    long long bad = 0;
    TSG_TRY(tsg_read_scalar_i64(ctx, seen + n, 0, &bad));
    tsg_free(ctx, seen);
    if (bad) { tsg_set_error("not a permutation of 0..%d", n - 1); return TSG_ERR_ARG; }
    return TSG_OK;
}

extern "C" int tsg_perm_invert(tsg_ctx *ctx, int n, const int *d_old_id, int *d_new_id) {
    if (n > 0) {
        hipLaunchKernelGGL(k_scatter_inverse, dim3(tsg_div_up(n, 256)), dim3(256), 0, ctx->stream, d_old_id, n, d_new_id);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

// ------------------------------------------------------------------ degree
__global__ void k_rowlen(const int *__restrict__ rowptr, int m, int *__restrict__ deg) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) deg[i] = rowptr[i + 1] - rowptr[i];
}
__global__ void k_mirror_deg(const int *__restrict__ rowptr, const int *__restrict__ col, int m, int *__restrict__ deg) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    for (int e = rowptr[i]; e < rowptr[i + 1]; e++)
        if (col[e] != i) atomicAdd(&deg[col[e]], 1);
}
__global__ void k_max_i32(const int *__restrict__ v, int n, int *__restrict__ out) {
    int mx = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) mx = max(mx, v[i]);
    mx = tsg_wave_max(mx);
    if ((threadIdx.x & 63) == 0) atomicMax(out, mx);
}
__global__ void k_deg_key(const int *__restrict__ deg, int n, int maxdeg, int descending, unsigned *__restrict__ key) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) key[i] = descending ? (unsigned)(maxdeg - deg[i]) : (unsigned)deg[i];
}

static int max_of(tsg_ctx *ctx, const int *v, int n, int *out) {
    void *scr;
    TSG_TRY(tsg_scratch(ctx, 64, &scr));
    TSG_HIP(hipMemsetAsync(scr, 0, 4, ctx->stream));
    if (n > 0) {
        hipLaunchKernelGGL(k_max_i32, dim3(std::min<unsigned>(1024, tsg_div_up(n, 256))), dim3(256), 0, ctx->stream, v, n,
                           (int *)scr);
        TSG_LAUNCH_CHECK();
    }
    long long m;
    TSG_TRY(tsg_read_scalar_i64(ctx, scr, 0, &m));
    *out = (int)m;
    return TSG_OK;
}

// ids sorted by degree (ascending or descending), ties by ascending id
static int sort_ids_by_degree(tsg_ctx *ctx, const int *deg, int n, int descending, int *ids) {
    int maxdeg = 0;
    TSG_TRY(max_of(ctx, deg, n, &maxdeg));
    unsigned *key;
    TSG_TRY(tsg_alloc_t(ctx, &key, (size_t)n + 1));
    if (n > 0) {
        hipLaunchKernelGGL(k_deg_key, dim3(tsg_div_up(n, 256)), dim3(256), 0, ctx->stream, deg, n, maxdeg, descending, key);
        hipLaunchKernelGGL(k_iota2, dim3(tsg_div_up(n, 256)), dim3(256), 0, ctx->stream, ids, (long long)n);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u32(ctx, key, ids, n, bits_of(maxdeg)));
    }
    tsg_free(ctx, key);
    return TSG_OK;
}

extern "C" int tsg_order_degree(tsg_ctx *ctx, const tsg_dcsr *a, int symmetric_header, int *d_old_id) {
    if (a->m != a->n) { tsg_set_error("degree order needs a square matrix"); return TSG_ERR_ARG; }
    // only row lengths (and, for the mirror count, column ids) are read:
    // a CSR straight from mmio_allinone, repeated entries included, is
    // ranked by its row lengths as given
    TSG_TRY(tsg_validate_csr_mode(ctx, a, 0));
    int *deg;
    TSG_TRY(tsg_alloc_t(ctx, &deg, (size_t)a->m + 1));
    const unsigned g = tsg_div_up(a->m, 256);
    if (a->m > 0) {
        hipLaunchKernelGGL(k_rowlen, dim3(g), dim3(256), 0, ctx->stream, a->rowptr, a->m, deg);
        if (symmetric_header) hipLaunchKernelGGL(k_mirror_deg, dim3(g), dim3(256), 0, ctx->stream, a->rowptr, a->colidx, a->m, deg);
        TSG_LAUNCH_CHECK();
    }
    TSG_TRY(sort_ids_by_degree(ctx, deg, a->m, 1, d_old_id));
    tsg_free(ctx, deg);
    return TSG_OK;
}

// ------------------------------------------------------------------ random
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
__global__ void k_rand_key(int n, unsigned long long seed, unsigned long long *__restrict__ key, int *__restrict__ ids) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    // high bits random, low bits the id: distinct keys, uniform order
    key[i] = (mix64(seed ^ mix64((unsigned long long)i)) & 0xffffffff00000000ULL) | (unsigned)i;
    ids[i] = i;
}

extern "C" int tsg_order_random(tsg_ctx *ctx, int n, unsigned long long seed, int *d_old_id) {
    unsigned long long *key;
    TSG_TRY(tsg_alloc_t(ctx, &key, (size_t)n + 1));
    if (n > 0) {
        hipLaunchKernelGGL(k_rand_key, dim3(tsg_div_up(n, 256)), dim3(256), 0, ctx->stream, n, seed, key, d_old_id);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u64(ctx, key, d_old_id, n, 64));
    }
    tsg_free(ctx, key);
    return TSG_OK;
}

// ------------------------------------------------------------------ symmetrised pattern
__global__ void k_fill_f64(double *v, long long n, double x) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = x;
}

template <bool FILL>
__global__ void __launch_bounds__(256)
k_union_rows(int m, const int *__restrict__ ap, const int *__restrict__ ac, const int *__restrict__ tp,
             const int *__restrict__ tc, int *__restrict__ cnt_or_ptr, int *__restrict__ gc) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    int a = ap[i], ae = ap[i + 1], b = tp[i], be = tp[i + 1];
    int out = FILL ? cnt_or_ptr[i] : 0;
    while (a < ae || b < be) {
        int v;
        if (b >= be || (a < ae && ac[a] < tc[b])) v = ac[a++];
        else if (a >= ae || tc[b] < ac[a]) v = tc[b++];
        else { v = ac[a]; a++; b++; }
        if (FILL) gc[out] = v;
        out++;
    }
    if (!FILL) cnt_or_ptr[i] = out;
}

// pattern(A + A^T), sorted rows, self loops kept
extern "C" int tsg_dcsr_symmetrize(tsg_ctx *ctx, const tsg_dcsr *a, tsg_dcsr *g) {
    memset(g, 0, sizeof(*g));
    if (a->m != a->n) { tsg_set_error("symmetrize needs a square matrix"); return TSG_ERR_ARG; }
    tsg_dcsr at;
    TSG_TRY(tsg_dcsr_transpose(ctx, a, &at));
    g->m = g->n = a->m;
    TSG_TRY(tsg_alloc_t(ctx, &g->rowptr, (size_t)a->m + 1));
    const unsigned gr = tsg_div_up(a->m, 256);
    if (a->m > 0) {
        hipLaunchKernelGGL(k_union_rows<false>, dim3(gr), dim3(256), 0, ctx->stream, a->m, a->rowptr, a->colidx,
                           at.rowptr, at.colidx, g->rowptr, (int *)nullptr);
        TSG_LAUNCH_CHECK();
    }
    long long nnz = 0;
    TSG_TRY(tsg_scan_i32(ctx, g->rowptr, g->rowptr, a->m, &nnz));
    g->nnz = (int)nnz;
    TSG_TRY(tsg_alloc_t(ctx, &g->colidx, (size_t)nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &g->val, (size_t)nnz + 1));
    if (a->m > 0 && nnz > 0) {
        hipLaunchKernelGGL(k_union_rows<true>, dim3(gr), dim3(256), 0, ctx->stream, a->m, a->rowptr, a->colidx,
                           at.rowptr, at.colidx, g->rowptr, g->colidx);
        hipLaunchKernelGGL(k_fill_f64, dim3(tsg_div_up(nnz, 256)), dim3(256), 0, ctx->stream, g->val, nnz, 1.0);
        TSG_LAUNCH_CHECK();
    }
    tsg_dcsr_free(ctx, &at);
    return TSG_OK;
}

// ------------------------------------------------------------------ RCM
// degree as SciPy's _node_degrees: row length + 1 when the diagonal is stored
__global__ void k_rcm_deg(const int *__restrict__ gp, const int *__restrict__ gc, int n, int *__restrict__ deg) {
    int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= n) return;
    int d = gp[v + 1] - gp[v];
    for (int e = gp[v]; e < gp[v + 1]; e++)
        if (gc[e] == v) { d++; break; }
    deg[v] = d;
}
__global__ void k_scatter_rank(const int *__restrict__ order, int n, int *__restrict__ rank, int *__restrict__ label) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    rank[order[r]] = r;
    label[order[r]] = r;
}
// label <- min label over the closed neighbourhood, then pointer jumping
// through the vertex that owns the label (labels only decrease).
__global__ void k_cc_prop(const int *__restrict__ gp, const int *__restrict__ gc, int n, int *label, int *changed) {
    int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= n) return;
    int l = label[v];
    const int l0 = l;
    for (int e = gp[v]; e < gp[v + 1]; e++) l = min(l, label[gc[e]]);
    if (l < l0) {
        atomicMin(&label[v], l);
        *changed = 1;
    }
}
__global__ void k_cc_jump(const int *__restrict__ order, int n, int *label) {
    int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= n) return;
    int l = label[v];
    for (int it = 0; it < 64; it++) {
        const int l2 = label[order[l]];
        if (l2 >= l) break;
        l = l2;
    }
    atomicMin(&label[v], l);
}
// seeds (component minima) in rank order: flag, scan, compact
__global__ void k_seed_flag(const int *__restrict__ order, const int *__restrict__ label, int n, int *__restrict__ flag) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) flag[r] = label[order[r]] == r;
}
__global__ void k_seed_compact(const int *__restrict__ order, const int *__restrict__ label, const int *__restrict__ pos,
                               int n, int *__restrict__ front, int *__restrict__ visited) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n && label[order[r]] == r) {
        front[pos[r]] = order[r];
        visited[order[r]] = 1;
    }
}
__global__ void k_claim(const int *__restrict__ gp, const int *__restrict__ gc, const int *__restrict__ front, int nf,
                        const int *__restrict__ visited, int *__restrict__ claim) {
    int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nf) return;
    const int p = front[q];
    for (int e = gp[p]; e < gp[p + 1]; e++) {
        const int u = gc[e];
        if (!visited[u]) atomicMin(&claim[u], q);
    }
}
template <bool EMIT>
__global__ void k_children(const int *__restrict__ gp, const int *__restrict__ gc, const int *__restrict__ front, int nf,
                           const int *__restrict__ visited, const int *__restrict__ claim, const int *__restrict__ deg,
                           int *__restrict__ cnt_or_off, unsigned long long *__restrict__ key, int *__restrict__ par) {
    int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nf) return;
    const int p = front[q];
    int k = EMIT ? cnt_or_off[q] : 0;
    for (int e = gp[p]; e < gp[p + 1]; e++) {
        const int u = gc[e];
        if (!visited[u] && claim[u] == q) {
            if (EMIT) {
                key[k] = ((unsigned long long)(unsigned)deg[u] << 32) | (unsigned)u;
                par[k] = q;
            }
            k++;
        }
    }
    if (!EMIT) cnt_or_off[q] = k;
}
__global__ void k_mark_children(const unsigned long long *__restrict__ key, int nc, int *__restrict__ visited) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nc) visited[(int)(key[i] & 0xffffffffu)] = 1;
}
__global__ void k_parent_key(const int *__restrict__ par, int nc, unsigned *__restrict__ k2) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nc) k2[i] = (unsigned)par[i];
}
__global__ void k_gather_children(const unsigned long long *__restrict__ key, const int *__restrict__ idx, int nc,
                                  int *__restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nc) out[i] = (int)(key[idx[i]] & 0xffffffffu);
}
__global__ void k_comp_key(const int *__restrict__ seq, const int *__restrict__ label, int n, unsigned *__restrict__ k) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) k[i] = (unsigned)label[seq[i]];
}
__global__ void k_reverse_gather(const int *__restrict__ seq, const int *__restrict__ idx, int n, int *__restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = seq[idx[n - 1 - i]];
}
__global__ void k_fill_i32(int *v, long long n, int x) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = x;
}

extern "C" int tsg_order_rcm(tsg_ctx *ctx, const tsg_dcsr *a, int *d_old_id, int *levels_out) {
    if (a->m != a->n) { tsg_set_error("RCM needs a square matrix"); return TSG_ERR_ARG; }
    const int n = a->m;
    if (n == 0) return TSG_OK;
    tsg_dcsr g;
    TSG_TRY(tsg_dcsr_symmetrize(ctx, a, &g));
    int *deg, *order, *rank, *label, *visited, *claim, *seq, *flag, *pos;
    TSG_TRY(tsg_alloc_t(ctx, &deg, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &order, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &rank, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &label, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &visited, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &claim, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &seq, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &flag, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &pos, (size_t)n + 1));
    hipStream_t s = ctx->stream;
    const unsigned gn = tsg_div_up(n, 256);
    hipLaunchKernelGGL(k_rcm_deg, dim3(gn), dim3(256), 0, s, g.rowptr, g.colidx, n, deg);
    TSG_LAUNCH_CHECK();
    TSG_TRY(sort_ids_by_degree(ctx, deg, n, 0, order));  // ascending (degree, id)
    hipLaunchKernelGGL(k_scatter_rank, dim3(gn), dim3(256), 0, s, order, n, rank, label);
    TSG_LAUNCH_CHECK();
    // connected components: label = lowest rank in the component
    void *scr;
    TSG_TRY(tsg_scratch(ctx, 64, &scr));
    for (int it = 0; it < 4 * n + 8; it++) {
        TSG_HIP(hipMemsetAsync(scr, 0, 4, s));
        hipLaunchKernelGGL(k_cc_prop, dim3(gn), dim3(256), 0, s, g.rowptr, g.colidx, n, label, (int *)scr);
        hipLaunchKernelGGL(k_cc_jump, dim3(gn), dim3(256), 0, s, order, n, label);
        TSG_LAUNCH_CHECK();
        long long ch = 0;
        TSG_TRY(tsg_read_scalar_i64(ctx, scr, 0, &ch));
        if (!ch) break;
    }
    // level 0: all component seeds, in rank order
    TSG_HIP(hipMemsetAsync(visited, 0, (size_t)n * 4, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(256), 0, s, claim, (long long)n, 0x7fffffff);
    hipLaunchKernelGGL(k_seed_flag, dim3(gn), dim3(256), 0, s, order, label, n, flag);
    TSG_LAUNCH_CHECK();
    long long nseed = 0;
    TSG_TRY(tsg_scan_i32(ctx, flag, pos, n, &nseed));
    hipLaunchKernelGGL(k_seed_compact, dim3(gn), dim3(256), 0, s, order, label, pos, n, seq, visited);
    TSG_LAUNCH_CHECK();
    long long done = nseed, fstart = 0, nf = nseed;
    int levels = 1;
    unsigned long long *key = nullptr;
    int *par = nullptr, *cnt = nullptr, *idx = nullptr;
    unsigned *k2 = nullptr;
    TSG_TRY(tsg_alloc_t(ctx, &key, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &par, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &cnt, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &idx, (size_t)n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &k2, (size_t)n + 1));
    const int nbits = bits_of(n);
    while (nf > 0 && done < n) {
        const int *front = seq + fstart;
        const unsigned gf = tsg_div_up(nf, 256);
        hipLaunchKernelGGL(k_claim, dim3(gf), dim3(256), 0, s, g.rowptr, g.colidx, front, (int)nf, visited, claim);
        hipLaunchKernelGGL((k_children<false>), dim3(gf), dim3(256), 0, s, g.rowptr, g.colidx, front, (int)nf, visited,
                           claim, deg, cnt, (unsigned long long *)nullptr, (int *)nullptr);
        TSG_LAUNCH_CHECK();
        long long nc = 0;
        TSG_TRY(tsg_scan_i32(ctx, cnt, cnt, nf, &nc));
        if (nc == 0) break;
        hipLaunchKernelGGL((k_children<true>), dim3(gf), dim3(256), 0, s, g.rowptr, g.colidx, front, (int)nf, visited,
                           claim, deg, cnt, key, par);
        TSG_LAUNCH_CHECK();
        const unsigned gc2 = tsg_div_up(nc, 256);
        hipLaunchKernelGGL(k_mark_children, dim3(gc2), dim3(256), 0, s, key, (int)nc, visited);
        // order children by (parent position, degree, id): sort by (degree,
        // id), then stably by parent position
        hipLaunchKernelGGL(k_iota2, dim3(gc2), dim3(256), 0, s, idx, nc);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u64(ctx, key, idx, nc, 32 + nbits));
        // after the sort, idx[i] is the emit slot of the i-th smallest key:
        // parent of the i-th smallest key is par[idx[i]]
        hipLaunchKernelGGL(k_gather_i32_r, dim3(gc2), dim3(256), 0, s, par, idx, (int)nc, (int *)k2);
        hipLaunchKernelGGL(k_iota2, dim3(gc2), dim3(256), 0, s, idx, nc);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u32(ctx, k2, idx, nc, bits_of(nf)));
        hipLaunchKernelGGL(k_gather_children, dim3(gc2), dim3(256), 0, s, key, idx, (int)nc, seq + fstart + nf);
        TSG_LAUNCH_CHECK();
        fstart += nf;
        nf = nc;
        done += nc;
        levels++;
    }
    if (done != n) { tsg_set_error("RCM: traversal covered %lld of %d vertices", done, n); return TSG_ERR_STATE; }
    // group the level-ordered sequence by component (stable), then reverse
    unsigned *ck;
    TSG_TRY(tsg_alloc_t(ctx, &ck, (size_t)n + 1));
    hipLaunchKernelGGL(k_comp_key, dim3(gn), dim3(256), 0, s, seq, label, n, ck);
    hipLaunchKernelGGL(k_iota2, dim3(gn), dim3(256), 0, s, idx, (long long)n);
    TSG_LAUNCH_CHECK();
    TSG_TRY(tsg_sort_pairs_u32(ctx, ck, idx, n, nbits));
    hipLaunchKernelGGL(k_reverse_gather, dim3(gn), dim3(256), 0, s, seq, idx, n, d_old_id);
    TSG_LAUNCH_CHECK();
    if (levels_out) *levels_out = levels;
    void *ps[] = {deg, order, rank, label, visited, claim, seq, flag, pos, key, par, cnt, idx, k2, ck};
    for (void *p : ps) tsg_free(ctx, p);
    tsg_dcsr_free(ctx, &g);
    return TSG_OK;
}

// ------------------------------------------------------------------ P A Q^T
__global__ void k_perm_counts(const int *__restrict__ rowptr, const int *__restrict__ row_old, int m, int *__restrict__ cnt) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < m) cnt[r] = rowptr[row_old[r] + 1] - rowptr[row_old[r]];
}
__global__ void k_perm_emit(const int *__restrict__ rowptr, const int *__restrict__ col, const int *__restrict__ row_old,
                            const int *__restrict__ col_new, int m, int n, const int *__restrict__ optr,
                            unsigned long long *__restrict__ key, int *__restrict__ src) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const int o = row_old[r];
    int w = optr[r];
    for (int e = rowptr[o]; e < rowptr[o + 1]; e++, w++) {
        key[w] = (unsigned long long)r * (unsigned long long)n + (unsigned)col_new[col[e]];
        src[w] = e;
    }
}
__global__ void k_perm_gather(const unsigned long long *__restrict__ key, const int *__restrict__ src,
                              const double *__restrict__ val, int nnz, int n, int *__restrict__ ocol, double *__restrict__ oval) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nnz) return;
    ocol[i] = (int)(key[i] % (unsigned long long)n);
    oval[i] = val[src[i]];
}

// out row r = a row row_old[r], columns relabelled c -> col_new[c], rows sorted.
extern "C" int tsg_dcsr_permute(tsg_ctx *ctx, const tsg_dcsr *a, const int *d_row_old, const int *d_col_new,
                                tsg_dcsr *out) {
    memset(out, 0, sizeof(*out));
    TSG_TRY(tsg_validate_csr(ctx, a));
    TSG_TRY(tsg_perm_validate(ctx, a->m, d_row_old));
    TSG_TRY(tsg_perm_validate(ctx, a->n, d_col_new));
    const int m = a->m, n = a->n, nnz = a->nnz;
    out->m = m;
    out->n = n;
    out->nnz = nnz;
    TSG_TRY(tsg_alloc_t(ctx, &out->rowptr, (size_t)m + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->colidx, (size_t)nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->val, (size_t)nnz + 1));
    const unsigned gm = tsg_div_up(m, 256);
    if (m > 0) {
        hipLaunchKernelGGL(k_perm_counts, dim3(gm), dim3(256), 0, ctx->stream, a->rowptr, d_row_old, m, out->rowptr);
        TSG_LAUNCH_CHECK();
    }
    long long tot = 0;
    TSG_TRY(tsg_scan_i32(ctx, out->rowptr, out->rowptr, m, &tot));
    if (tot != nnz) { tsg_set_error("permute: row permutation is not a bijection"); return TSG_ERR_ARG; }
    if (nnz > 0) {
        unsigned long long *key;
        int *src;
        TSG_TRY(tsg_alloc_t(ctx, &key, (size_t)nnz));
        TSG_TRY(tsg_alloc_t(ctx, &src, (size_t)nnz));
        hipLaunchKernelGGL(k_perm_emit, dim3(gm), dim3(256), 0, ctx->stream, a->rowptr, a->colidx, d_row_old, d_col_new,
                           m, n, out->rowptr, key, src);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u64(ctx, key, src, nnz, bits_of((long long)m * n)));
        hipLaunchKernelGGL(k_perm_gather, dim3(tsg_div_up(nnz, 256)), dim3(256), 0, ctx->stream, key, src, a->val, nnz, n,
                           out->colidx, out->val);
        TSG_LAUNCH_CHECK();
        tsg_free(ctx, key);
        tsg_free(ctx, src);
    }
    return TSG_OK;
}
