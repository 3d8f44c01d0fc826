// spgemm.hip -- the TileSpGEMM pipeline on gfx950 (wave64, LDS, FP64 FMA).
//
// Step 1 (tile-level symbolic, src/tilespgemm-cuda.h:257-369 and the nsparse
//   hash path src/spgemm_nsparse_kernel.h:138-1511): for each A tile-row I the
//   set of C tile columns  U_{K in A(I,:)} B(K,:)  -- one wave per tile-row,
//   the (K, J) candidates flattened across the 64 lanes and OR-ed into a
//   windowed LDS bitmap; count pass, device scan, then a fill pass emitting
//   the sorted tile columns.  Covers every tilen (no SPA/hash split).
// Step 2 (16x16 bitmask symbolic, src/tilespgemm-cuda.h:371-755): one C tile
//   per 16-lane quarter-wave, lane r owns tile row r.  The matched K tiles
//   come from a binary-search intersection of A's tile-row list and B's
//   tile-column list, 16 keys per round, ordered with a 64-bit ballot.
//   Row mask r = OR over pairs, over k in Amask[r], of Bmask[k]; popcount,
//   in-group scan -> u8 row offsets and tile nnz.
// Step 3 (numeric, src/tilespgemm-cuda.h:1505-2608): same tile/lane mapping.
//   Lane r walks its A row in ascending k and each selected B row in
//   ascending column, accumulating with fma into an LDS accumulator that
//   only it touches: a 16-wide dense row (tiles with many nonzeros) or the
//   compact row segment of the C tile (sparse tiles).  No atomics: every C
//   entry receives its products in ascending inner index k, the order of the
//   serial SPA reference, so results are deterministic and bit-identical to
//   it.  (The reference uses FP64 atomicAdd, src/tilespgemm-cuda.h:1670.)
#include <string.h>

#include "tsg_internal.h"

#define S1_WAVES 4
#define S23_WAVES 4
#define TILES_PER_WAVE 4

extern "C" void tsg_tune_default(tsg_tune *t) {
    memset(t, 0, sizeof(*t));
    t->dense_acc_min_nnz = 64;
    t->s1_window_words = 256;
}

// ------------------------------------------------------------------ step 1
template <bool FILL, int BMW>
__global__ void __launch_bounds__(64 * S1_WAVES)
k_step1(const int *__restrict__ Aptr, const int *__restrict__ Acol, int tilem,
        const int *__restrict__ Bptr, const int *__restrict__ Bcol,
        int *__restrict__ Cptr_or_cnt, int *__restrict__ Ccol, int *__restrict__ Crow) {
    __shared__ unsigned s_bm[S1_WAVES][BMW];
    __shared__ int s_pre[S1_WAVES][64];
    __shared__ int s_b0[S1_WAVES][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = blockIdx.x * S1_WAVES + wave;
    if (I >= tilem) return;
    unsigned *bm = s_bm[wave];
    int *pre = s_pre[wave], *pb0 = s_b0[wave];
    const int a0 = Aptr[I], a1 = Aptr[I + 1];

    // range of candidate tile columns: B rows are sorted
    int jmin = 0x7fffffff, jmax = -1;
    for (int p = a0 + lane; p < a1; p += 64) {
        const int K = Acol[p];
        const int b0 = Bptr[K], b1 = Bptr[K + 1];
        if (b1 > b0) {
            jmin = min(jmin, Bcol[b0]);
            jmax = max(jmax, Bcol[b1 - 1]);
        }
    }
    jmin = tsg_wave_min(jmin);
    jmax = tsg_wave_max(jmax);

    int total = 0;
    const int base = FILL ? Cptr_or_cnt[I] : 0;
    for (int w0 = jmin; w0 <= jmax; w0 += 32 * BMW) {
        const int nw = min(BMW, ((jmax - w0) >> 5) + 1);
        for (int k = lane; k < nw; k += 64) bm[k] = 0u;
        tsg_wave_sync();
        for (int c0 = a0; c0 < a1; c0 += 64) {
            const int p = c0 + lane;
            int b0 = 0, len = 0;
            if (p < a1) {
                const int K = Acol[p];
                b0 = Bptr[K];
                len = Bptr[K + 1] - b0;
            }
            const int incl = tsg_wave_incl_scan(len);
            const int tot = __shfl(incl, 63, 64);
            pre[lane] = incl - len;
            pb0[lane] = b0;
            tsg_wave_sync();
            for (int c = lane; c < tot; c += 64) {
                int lo = 0;
#pragma unroll
                for (int st = 32; st > 0; st >>= 1)
                    if (pre[lo + st] <= c) lo += st;
                const int d = Bcol[pb0[lo] + (c - pre[lo])] - w0;
                if (d >= 0 && d < 32 * BMW) atomicOr(&bm[d >> 5], 1u << (d & 31));
            }
            tsg_wave_sync();
        }
        if (!FILL) {
            int cnt = 0;
            for (int k = lane; k < nw; k += 64) cnt += __popc(bm[k]);
            total += tsg_wave_sum(cnt);
        } else {
            const int chunk = (nw + 63) >> 6;
            const int k0 = lane * chunk, k1 = min(nw, k0 + chunk);
            int cnt = 0;
            for (int k = k0; k < k1; k++) cnt += __popc(bm[k]);
            int pos = base + total + tsg_wave_incl_scan(cnt) - cnt;
            for (int k = k0; k < k1; k++) {
                unsigned x = bm[k];
                while (x) {
                    const int b = __ffs(x) - 1;
                    x &= x - 1;
                    Ccol[pos] = w0 + k * 32 + b;
                    Crow[pos] = I;
                    pos++;
                }
            }
            total += tsg_wave_sum(cnt);
        }
        tsg_wave_sync();
    }
    if (!FILL && lane == 0) Cptr_or_cnt[I] = total;
}

// ------------------------------------------------------------------ pairs
// Operands of steps 2 and 3.
struct TileOperands {
    const int *Aptr, *Acol, *Annz;
    const unsigned char *Aptr16;
    const unsigned short *Amask;
    const double *Aval;
    const int *Bcscptr, *Bcscrow, *Bcscid, *Bnnz;
    const unsigned char *Bptr16;
    const unsigned short *Bmask;
    const double *Bval;
};

// Largest index p in [0, len) with list[p] == key, or -1.
__device__ __forceinline__ int find_sorted(const int *__restrict__ list, int len, int key) {
    int lo = 0, hi = len - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = list[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

// Calls body(ta, tb) for every matched pair (A tile ta, B tile tb) of C tile
// (I, J), in ascending inner tile index K, uniformly across the 16 lanes of
// the calling group.  `g` is the group (0..3) inside the wave, `r` the lane
// inside the group.  Must be entered by all 16 lanes of the group.
template <typename Body>
__device__ __forceinline__ void for_each_pair(const TileOperands &op, int I, int J, int g, int r, Body body) {
    const int a0 = op.Aptr[I], la = op.Aptr[I + 1] - a0;
    const int b0 = op.Bcscptr[J], lb = op.Bcscptr[J + 1] - b0;
    const bool a_short = la <= lb;
    const int *S = a_short ? op.Acol + a0 : op.Bcscrow + b0;
    const int *L = a_short ? op.Bcscrow + b0 : op.Acol + a0;
    const int ls = a_short ? la : lb, ll = a_short ? lb : la;
    if (ls == 0) return;
    const int src_base = g * 16;
    for (int base = 0; base < ls; base += 16) {
        const int idx = base + r;
        int pos = -1;
        if (idx < ls) pos = find_sorted(L, ll, S[idx]);
        const unsigned long long bal = __ballot(pos >= 0);
        unsigned m16 = (unsigned)(bal >> src_base) & 0xffffu;
        while (m16) {
            const int j = __ffs(m16) - 1;
            m16 &= m16 - 1;
            const int sidx = base + j;
            const int lpos = __shfl(pos, src_base + j, 64);
            const int apos = a_short ? sidx : lpos;
            const int bpos = a_short ? lpos : sidx;
            body(a0 + apos, op.Bcscid[b0 + bpos]);
        }
    }
}

// ------------------------------------------------------------------ step 2
__global__ void __launch_bounds__(64 * S23_WAVES)
k_step2(TileOperands op, const int *__restrict__ Crow, const int *__restrict__ Ccol, int numtileC,
        unsigned short *__restrict__ Cmask, unsigned char *__restrict__ Cptr16, int *__restrict__ Cnnz,
        int dense_min, unsigned long long *__restrict__ bins) {
    __shared__ unsigned short s_bm[S23_WAVES][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const unsigned lb = tsg_xcd_block(blockIdx.x, gridDim.x);
    const long long t = ((long long)lb * S23_WAVES + wave) * TILES_PER_WAVE + g;
    if (t >= numtileC) return;  // whole groups exit together
    unsigned short *bmg = &s_bm[wave][g * 16];
    const int I = Crow[t], J = Ccol[t];
    unsigned orow = 0;
    for_each_pair(op, I, J, g, r, [&](int ta, int tb) {
        bmg[r] = op.Bmask[(long long)tb * 16 + r];
        tsg_wave_sync();
        unsigned am = op.Amask[(long long)ta * 16 + r];
        while (am) {
            const int b = 31 - __clz(am);
            am &= ~(1u << b);
            orow |= bmg[15 - b];
        }
        tsg_wave_sync();
    });
    const int cnt = __popc(orow);
    const int incl = tsg_group16_incl_scan(cnt);
    const int tot = __shfl(incl, g * 16 + 15, 64);
    Cmask[t * 16 + r] = (unsigned short)orow;
    Cptr16[t * 16 + r] = (unsigned char)(incl - cnt);
    if (r == 0) {
        Cnnz[t] = tot;
        atomicAdd(&bins[tot == 0 ? 0 : (tot >= dense_min ? 2 : 1)], 1ull);
    }
}

// ------------------------------------------------------------------ step 3
#define ACC_STRIDE 17  // doubles per lane row (padding against bank conflicts)

__global__ void __launch_bounds__(64 * S23_WAVES)
k_step3(TileOperands op, const int *__restrict__ Crow, const int *__restrict__ Ccol, int numtileC,
        const unsigned short *__restrict__ Cmask, const unsigned char *__restrict__ Cptr16,
        const int *__restrict__ Cnnz, unsigned char *__restrict__ Ccol8, double *__restrict__ Cval,
        int dense_min) {
    __shared__ double s_acc[S23_WAVES][64 * ACC_STRIDE];
    __shared__ unsigned short s_bm[S23_WAVES][64];
    __shared__ unsigned char s_bp[S23_WAVES][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const unsigned lblk = tsg_xcd_block(blockIdx.x, gridDim.x);
    const long long t = ((long long)lblk * S23_WAVES + wave) * TILES_PER_WAVE + g;
    if (t >= numtileC) return;
    const int coff = Cnnz[t], cn = Cnnz[t + 1] - coff;
    if (cn == 0) return;  // structurally empty tile (step-1 phantom)
    const unsigned cm = Cmask[t * 16 + r];
    const int cp = Cptr16[t * 16 + r];
    const int rowlen = __popc(cm);
    const bool dense = cn >= dense_min;
    double *accg = &s_acc[wave][g * 16 * ACC_STRIDE];  // 272 doubles per group
    double *accr = accg + r * ACC_STRIDE;               // dense: lane-private row
    if (dense) {
#pragma unroll
        for (int c = 0; c < 16; c++) accr[c] = 0.0;
    } else {
        for (int j = 0; j < rowlen; j++) accg[cp + j] = 0.0;
    }
    unsigned short *bmg = &s_bm[wave][g * 16];
    unsigned char *bpg = &s_bp[wave][g * 16];
    const int I = Crow[t], J = Ccol[t];
    for_each_pair(op, I, J, g, r, [&](int ta, int tb) {
        bmg[r] = op.Bmask[(long long)tb * 16 + r];
        bpg[r] = op.Bptr16[(long long)tb * 16 + r];
        tsg_wave_sync();
        const double *bval = op.Bval + op.Bnnz[tb];
        unsigned am = op.Amask[(long long)ta * 16 + r];
        const double *aval = op.Aval + op.Annz[ta] + op.Aptr16[(long long)ta * 16 + r];
        int ia = 0;
        while (am) {
            const int b = 31 - __clz(am);  // highest bit = smallest column k
            am &= ~(1u << b);
            const int k = 15 - b;
            const double a = aval[ia++];
            unsigned bm = bmg[k];
            const double *brow = bval + bpg[k];
            int jb = 0;
            while (bm) {
                const int bb = 31 - __clz(bm);
                bm &= ~(1u << bb);
                const double bv = brow[jb++];
                double *slot = dense ? &accr[15 - bb] : &accg[cp + __popc(cm >> (bb + 1))];
                *slot = __builtin_fma(a, bv, *slot);
            }
        }
        tsg_wave_sync();
    });
    // write out row r of the tile, ascending columns
    double *outv = Cval + coff + cp;
    unsigned char *outc = Ccol8 + coff + cp;
    unsigned x = cm;
    for (int j = 0; j < rowlen; j++) {
        const int bb = 31 - __clz(x);
        x &= ~(1u << bb);
        const int c = 15 - bb;
        outv[j] = dense ? accr[c] : accg[cp + j];
        outc[j] = (unsigned char)c;
    }
}

// ------------------------------------------------------------------ host
static int ev_ms(hipEvent_t a, hipEvent_t b, double *ms) {
    float f = 0.f;
    TSG_HIP(hipEventElapsedTime(&f, a, b));
    *ms = f;
    return TSG_OK;
}

extern "C" int tsg_spgemm(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, const tsg_tune *tune_in,
                          tsg_dtiles *C, tsg_stats *stats) {
    memset(C, 0, sizeof(*C));
    if (!A || !B || A->n != B->m) {
        tsg_set_error("tsg_spgemm: inner dimensions differ (%d vs %d)", A ? A->n : -1, B ? B->m : -1);
        return TSG_ERR_ARG;
    }
    if (!B->csc_ptr) { tsg_set_error("tsg_spgemm: B needs its CSC tile index (tsg_tiles_build_csc)"); return TSG_ERR_STATE; }
    tsg_tune tune;
    if (tune_in) tune = *tune_in; else tsg_tune_default(&tune);
    if (tune.dense_acc_min_nnz < 1) tune.dense_acc_min_nnz = 1;

    hipEvent_t ev[5];
    const bool timed = stats != nullptr;
    if (timed) {
        memset(stats, 0, sizeof(*stats));
        for (auto &e : ev) TSG_HIP(hipEventCreate(&e));
        TSG_HIP(hipEventRecord(ev[0], ctx->stream));
    }
    C->m = A->m;
    C->n = B->n;
    C->tilem = A->tilem;
    C->tilen = B->tilen;
    // ---- step 1
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_ptr, (size_t)C->tilem + 1));
    if (C->tilem > 0) {
        hipLaunchKernelGGL((k_step1<false, 256>), dim3(tsg_div_up(C->tilem, S1_WAVES)), dim3(64 * S1_WAVES), 0,
                           ctx->stream, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr, B->tile_colidx,
                           C->tile_ptr, (int *)nullptr, (int *)nullptr);
        TSG_LAUNCH_CHECK();
    }
    long long numtileC = 0;
    TSG_TRY(tsg_scan_i32(ctx, C->tile_ptr, C->tile_ptr, C->tilem, &numtileC));
    C->numtile = (int)numtileC;
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_colidx, (size_t)numtileC + 1));
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_rowidx, (size_t)numtileC + 1));
    if (numtileC > 0) {
        hipLaunchKernelGGL((k_step1<true, 256>), dim3(tsg_div_up(C->tilem, S1_WAVES)), dim3(64 * S1_WAVES), 0,
                           ctx->stream, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr, B->tile_colidx,
                           C->tile_ptr, C->tile_colidx, C->tile_rowidx);
        TSG_LAUNCH_CHECK();
    }
    if (timed) TSG_HIP(hipEventRecord(ev[1], ctx->stream));
    // ---- step 2
    TSG_TRY(tsg_alloc_t(ctx, &C->mask, (size_t)numtileC * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &C->csr_ptr, (size_t)numtileC * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_nnz, (size_t)numtileC + 1));
    void *scr;
    TSG_TRY(tsg_alloc(ctx, &scr, 64));
    unsigned long long *bins = (unsigned long long *)scr;
    TSG_HIP(hipMemsetAsync(bins, 0, 4 * sizeof(unsigned long long), ctx->stream));
    TileOperands op;
    op.Aptr = A->tile_ptr; op.Acol = A->tile_colidx; op.Annz = A->tile_nnz; op.Aptr16 = A->csr_ptr;
    op.Amask = A->mask; op.Aval = A->val;
    op.Bcscptr = B->csc_ptr; op.Bcscrow = B->csc_rowidx; op.Bcscid = B->csc_tileid; op.Bnnz = B->tile_nnz;
    op.Bptr16 = B->csr_ptr; op.Bmask = B->mask; op.Bval = B->val;
    const unsigned tiles_per_block = S23_WAVES * TILES_PER_WAVE;
    const unsigned nblk = tsg_div_up(numtileC, tiles_per_block);
    if (numtileC > 0) {
        hipLaunchKernelGGL(k_step2, dim3(nblk), dim3(64 * S23_WAVES), 0, ctx->stream, op, C->tile_rowidx,
                           C->tile_colidx, C->numtile, C->mask, C->csr_ptr, C->tile_nnz, tune.dense_acc_min_nnz, bins);
        TSG_LAUNCH_CHECK();
    }
    long long nnzC = 0;
    TSG_TRY(tsg_scan_i32(ctx, C->tile_nnz, C->tile_nnz, numtileC, &nnzC));
    C->nnz = (int)nnzC;
    if (timed) TSG_HIP(hipEventRecord(ev[2], ctx->stream));
    // ---- step 3
    TSG_TRY(tsg_alloc_t(ctx, &C->csr_col, (size_t)nnzC + 1));
    TSG_TRY(tsg_alloc_t(ctx, &C->val, (size_t)nnzC + 1));
    if (nnzC > 0) {
        hipLaunchKernelGGL(k_step3, dim3(nblk), dim3(64 * S23_WAVES), 0, ctx->stream, op, C->tile_rowidx,
                           C->tile_colidx, C->numtile, C->mask, C->csr_ptr, C->tile_nnz, C->csr_col, C->val,
                           tune.dense_acc_min_nnz);
        TSG_LAUNCH_CHECK();
    }
    if (timed) {
        TSG_HIP(hipEventRecord(ev[3], ctx->stream));
        TSG_HIP(hipEventSynchronize(ev[3]));
        TSG_TRY(ev_ms(ev[0], ev[1], &stats->t_step1_ms));
        TSG_TRY(ev_ms(ev[1], ev[2], &stats->t_step2_ms));
        TSG_TRY(ev_ms(ev[2], ev[3], &stats->t_step3_ms));
        TSG_TRY(ev_ms(ev[0], ev[3], &stats->t_total_ms));
        unsigned long long hb[4];
        TSG_HIP(hipMemcpyAsync(hb, bins, sizeof(hb), hipMemcpyDeviceToHost, ctx->stream));
        TSG_HIP(hipStreamSynchronize(ctx->stream));
        stats->tiles_empty = (long long)hb[0];
        stats->tiles_sparse_acc = (long long)hb[1];
        stats->tiles_dense_acc = (long long)hb[2];
        stats->numtileC = numtileC;
        stats->nnzC = nnzC;
        for (auto &e : ev) hipEventDestroy(e);
    }
    tsg_free(ctx, scr);
    return TSG_OK;
}
