// convert.hip -- CSR <-> 16x16 tile conversions, CSC tile index, stable CSR
// transpose and the intermediate-product count, all on the device.
//
// csr2tile follows the output contract of csr2tile_row_major
// (src/csr2tile.h:196-259): tiles sorted by tile column inside each
// tile-row, per-tile exclusive nnz offsets, per-row u8 offsets inside a
// tile, MSB-first u16 row masks, values row-major inside a tile with
// ascending columns.  The reference builds it on the host with per-thread
// dense flag arrays of tilen entries; here each tile-row is one wave64 with a
// windowed LDS bitmap over the tile columns it touches, followed by one
// thread per matrix row walking its (sorted) entries against the tile list.
#include <string.h>

#include "tsg_internal.h"

#define WAVES_PER_BLOCK 4
#define BM_WORDS 256  // LDS bitmap window per wave: 256 x 32 = 8192 tile columns

// ---------------------------------------------------------------- tile-row bitmaps
// Tile columns touched by rows [16I, 16I+16): entry e -> colidx[e] >> 4.
// Rows are sorted, so the window range comes from first/last entries.
template <bool FILL>
__global__ void __launch_bounds__(64 * WAVES_PER_BLOCK)
k_c2t_tiles(const int *__restrict__ rowptr, const int *__restrict__ colidx, int m, int tilem,
            int *__restrict__ cnt_or_ptr, int *__restrict__ out_col, int *__restrict__ out_row) {
    __shared__ unsigned bm_all[WAVES_PER_BLOCK][BM_WORDS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = blockIdx.x * WAVES_PER_BLOCK + wave;
    if (I >= tilem) return;
    unsigned *bm = bm_all[wave];
    const int r0 = I * TSG_TILE, r1 = min(r0 + TSG_TILE, m);
    const int e0 = rowptr[r0], e1 = rowptr[r1];
    int jmin = 0x7fffffff, jmax = -1;
    if (lane < r1 - r0) {
        int a = rowptr[r0 + lane], b = rowptr[r0 + lane + 1];
        if (b > a) { jmin = colidx[a] >> 4; jmax = colidx[b - 1] >> 4; }
    }
    jmin = tsg_wave_min(jmin);
    jmax = tsg_wave_max(jmax);
    int total = 0;
    int base = FILL ? cnt_or_ptr[I] : 0;
    for (int w0 = jmin; w0 <= jmax; w0 += 32 * BM_WORDS) {
        const int nw = min(BM_WORDS, ((jmax - w0) >> 5) + 1);
        for (int k = lane; k < nw; k += 64) bm[k] = 0u;
        tsg_wave_sync();
        for (int e = e0 + lane; e < e1; e += 64) {
            int d = (colidx[e] >> 4) - w0;
            if (d >= 0 && d < 32 * BM_WORDS) atomicOr(&bm[d >> 5], 1u << (d & 31));
        }
        tsg_wave_sync();
        if (!FILL) {
            int c = 0;
            for (int k = lane; k < nw; k += 64) c += __popc(bm[k]);
            total += tsg_wave_sum(c);
        } else {
            const int chunk = (nw + 63) >> 6;
            const int k0 = lane * chunk, k1 = min(nw, k0 + chunk);
            int c = 0;
            for (int k = k0; k < k1; k++) c += __popc(bm[k]);
            int excl = tsg_wave_incl_scan(c) - c;
            int pos = base + total + excl;
            for (int k = k0; k < k1; k++) {
                unsigned x = bm[k];
                while (x) {
                    int b = __ffs(x) - 1;
                    x &= x - 1;
                    out_col[pos] = w0 + k * 32 + b;
                    out_row[pos] = I;
                    pos++;
                }
            }
            total += tsg_wave_sum(c);
        }
        tsg_wave_sync();
    }
    if (!FILL && lane == 0) cnt_or_ptr[I] = total;
}

// One thread per matrix row: walk its sorted entries against the sorted
// tile-column list of its tile-row.  PASS 0 writes per-(tile,row) counts and
// masks; PASS 1 writes values and in-tile columns at their final offsets.
template <int PASS>
__global__ void __launch_bounds__(256)
k_c2t_rowwalk(const int *__restrict__ rowptr, const int *__restrict__ colidx, const double *__restrict__ val,
              int m, const int *__restrict__ tile_ptr, const int *__restrict__ tile_colidx,
              unsigned char *__restrict__ cnt16, unsigned short *__restrict__ mask,
              const int *__restrict__ tile_nnz, double *__restrict__ tval, unsigned char *__restrict__ tcol) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int I = i >> 4, r = i & 15;
    int t = tile_ptr[I];
    const int tend = tile_ptr[I + 1] - 1;
    int curJ = -1, n = 0, base = 0;
    unsigned mk = 0;
    const int e1 = rowptr[i + 1];
    for (int e = rowptr[i]; e < e1; e++) {
        const int col = colidx[e];
        const int J = col >> 4, c = col & 15;
        if (J != curJ) {
            if (PASS == 0 && curJ >= 0) {
                cnt16[(long long)t * 16 + r] = (unsigned char)n;
                mask[(long long)t * 16 + r] = (unsigned short)mk;
            }
            while (t < tend && tile_colidx[t] < J) t++;  // input validated sorted: always finds J
            curJ = J;
            n = 0;
            mk = 0;
            if (PASS == 1) base = tile_nnz[t] + cnt16[(long long)t * 16 + r];
        }
        if (PASS == 0) {
            mk |= 0x8000u >> c;
        } else {
            tval[base + n] = val[e];
            tcol[base + n] = (unsigned char)c;
        }
        n++;
    }
    if (PASS == 0 && curJ >= 0) {
        cnt16[(long long)t * 16 + r] = (unsigned char)n;
        mask[(long long)t * 16 + r] = (unsigned short)mk;
    }
}

// One thread per tile: per-row counts -> exclusive u8 offsets, tile nnz.
__global__ void __launch_bounds__(256)
k_c2t_finalize(int numtile, unsigned char *__restrict__ ptr16, int *__restrict__ tnnz) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= numtile) return;
    uint4 v = *reinterpret_cast<const uint4 *>(ptr16 + (long long)t * 16);
    unsigned w[4] = {v.x, v.y, v.z, v.w};
    unsigned o[4];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        unsigned out = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            unsigned cnt = (w[q] >> (8 * b)) & 0xffu;
            out |= (unsigned)s << (8 * b);
            s += cnt;
        }
        o[q] = out;
    }
    *reinterpret_cast<uint4 *>(ptr16 + (long long)t * 16) = make_uint4(o[0], o[1], o[2], o[3]);
    tnnz[t] = s;
}

// Input contract of the conversion: row pointers non-decreasing from 0 to
// nnz, columns in [0, n) and (when `sorted`) strictly increasing inside each
// row.  Any violation sets a flag instead of letting a later kernel index
// out of range.
__global__ void __launch_bounds__(256)
k_validate_csr(const int *__restrict__ rowptr, const int *__restrict__ colidx, int m, int n, int nnz,
               int sorted, int *__restrict__ bad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i > m) return;
    const int s = rowptr[i];
    if (i == 0 && s != 0) atomicOr(bad, 1);
    if (i == m) {
        if (s != nnz) atomicOr(bad, 1);
        return;
    }
    const int e = rowptr[i + 1];
    if (s < 0 || e < s || e > nnz) { atomicOr(bad, 1); return; }
    int prev = -1;
    for (int k = s; k < e; k++) {
        const int c = colidx[k];
        if ((sorted && c <= prev) || c < 0 || c >= n) { atomicOr(bad, 2); return; }
        prev = c;
    }
}

int tsg_validate_csr_mode(tsg_ctx *ctx, const tsg_dcsr *a, int sorted) {
    void *scr;
    TSG_TRY(tsg_scratch(ctx, 64, &scr));
    TSG_HIP(hipMemsetAsync(scr, 0, 4, ctx->stream));
    hipLaunchKernelGGL(k_validate_csr, dim3(tsg_div_up((long long)a->m + 1, 256)), dim3(256), 0, ctx->stream,
                       a->rowptr, a->colidx, a->m, a->n, a->nnz, sorted, (int *)scr);
    TSG_LAUNCH_CHECK();
    long long bad = 0;
    TSG_TRY(tsg_read_scalar_i64(ctx, scr, 0, &bad));
    if (bad) {
        tsg_set_error(bad & 1 ? "CSR row pointers inconsistent with nnz"
                      : sorted ? "CSR columns out of range or not strictly increasing inside a row"
                               : "CSR columns out of range");
        return TSG_ERR_ARG;
    }
    return TSG_OK;
}

int tsg_validate_csr(tsg_ctx *ctx, const tsg_dcsr *a) { return tsg_validate_csr_mode(ctx, a, 1); }

extern "C" int tsg_csr2tile(tsg_ctx *ctx, const tsg_dcsr *a, tsg_dtiles *out) {
    memset(out, 0, sizeof(*out));
    if (!a || a->m < 0 || a->n < 0 || a->nnz < 0) { tsg_set_error("tsg_csr2tile: bad matrix"); return TSG_ERR_ARG; }
    TSG_TRY(tsg_validate_csr(ctx, a));
    tsg_dtiles &T = *out;
    T.m = a->m;
    T.n = a->n;
    T.nnz = a->nnz;
    T.tilem = (a->m + 15) / 16;
    T.tilen = (a->n + 15) / 16;
    TSG_TRY(tsg_alloc_t(ctx, &T.tile_ptr, (size_t)T.tilem + 1));
    if (T.tilem > 0) {
        hipLaunchKernelGGL(k_c2t_tiles<false>, dim3(tsg_div_up(T.tilem, WAVES_PER_BLOCK)), dim3(64 * WAVES_PER_BLOCK),
                           0, ctx->stream, a->rowptr, a->colidx, a->m, T.tilem, T.tile_ptr, (int *)nullptr,
                           (int *)nullptr);
        TSG_LAUNCH_CHECK();
    }
    long long numtile = 0;
    TSG_TRY(tsg_scan_i32(ctx, T.tile_ptr, T.tile_ptr, T.tilem, &numtile));
    T.numtile = (int)numtile;
    TSG_TRY(tsg_alloc_t(ctx, &T.tile_colidx, (size_t)numtile + 1));
    TSG_TRY(tsg_alloc_t(ctx, &T.tile_rowidx, (size_t)numtile + 1));
    TSG_TRY(tsg_alloc_t(ctx, &T.tile_nnz, (size_t)numtile + 1));
    TSG_TRY(tsg_alloc_t(ctx, &T.csr_ptr, (size_t)numtile * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &T.mask, (size_t)numtile * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &T.csr_col, (size_t)a->nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &T.val, (size_t)a->nnz + 1));
    TSG_HIP(hipMemsetAsync(T.csr_ptr, 0, (size_t)numtile * 16 + 16, ctx->stream));
    TSG_HIP(hipMemsetAsync(T.mask, 0, ((size_t)numtile * 16 + 16) * 2, ctx->stream));
    if (numtile > 0) {
        hipLaunchKernelGGL(k_c2t_tiles<true>, dim3(tsg_div_up(T.tilem, WAVES_PER_BLOCK)), dim3(64 * WAVES_PER_BLOCK),
                           0, ctx->stream, a->rowptr, a->colidx, a->m, T.tilem, T.tile_ptr, T.tile_colidx,
                           T.tile_rowidx);
        TSG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_c2t_rowwalk<0>, dim3(tsg_div_up(a->m, 256)), dim3(256), 0, ctx->stream, a->rowptr,
                           a->colidx, a->val, a->m, T.tile_ptr, T.tile_colidx, T.csr_ptr, T.mask, (const int *)nullptr,
                           (double *)nullptr, (unsigned char *)nullptr);
        TSG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_c2t_finalize, dim3(tsg_div_up(numtile, 256)), dim3(256), 0, ctx->stream, T.numtile,
                           T.csr_ptr, T.tile_nnz);
        TSG_LAUNCH_CHECK();
    }
    long long nnz = 0;
    TSG_TRY(tsg_scan_i32(ctx, T.tile_nnz, T.tile_nnz, numtile, &nnz));
    if (nnz != a->nnz) {
        tsg_set_error("tsg_csr2tile: tile nnz %lld != csr nnz %d (rows not sorted/unique?)", nnz, a->nnz);
        return TSG_ERR_ARG;
    }
    if (numtile > 0 && a->m > 0) {
        hipLaunchKernelGGL(k_c2t_rowwalk<1>, dim3(tsg_div_up(a->m, 256)), dim3(256), 0, ctx->stream, a->rowptr,
                           a->colidx, a->val, a->m, T.tile_ptr, T.tile_colidx, T.csr_ptr, T.mask, T.tile_nnz, T.val,
                           T.csr_col);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

// ---------------------------------------------------------------- CSC tile index
__global__ void k_iota(int *v, long long n) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (int)i;
}
__global__ void k_count_keys(const int *__restrict__ k, long long n, int *__restrict__ cnt) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) atomicAdd(&cnt[k[i]], 1);
}
__global__ void k_gather_i32(const int *__restrict__ src, const int *__restrict__ idx, long long n, int *__restrict__ dst) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_scatter_i32(const int *__restrict__ src, const int *__restrict__ idx, long long n, int *__restrict__ dst) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}
__global__ void k_gather_f64(const double *__restrict__ src, const int *__restrict__ idx, long long n, double *__restrict__ dst) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_u32_from_i32(const int *__restrict__ src, long long n, unsigned *__restrict__ dst) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (unsigned)src[i];
}

static int bits_for(long long v) {
    int b = 1;
    while ((1ll << b) <= v) b++;
    return b;
}

extern "C" int tsg_tiles_build_csc(tsg_ctx *ctx, tsg_dtiles *t) {
    if (t->csc_ptr) return TSG_OK;
    const long long nt = t->numtile;
    TSG_TRY(tsg_alloc_t(ctx, &t->csc_ptr, (size_t)t->tilen + 1));
    TSG_TRY(tsg_alloc_t(ctx, &t->csc_rowidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &t->csc_tileid, (size_t)nt + 1));
    TSG_HIP(hipMemsetAsync(t->csc_ptr, 0, ((size_t)t->tilen + 1) * 4, ctx->stream));
    if (nt > 0) {
        unsigned *keys;
        TSG_TRY(tsg_alloc_t(ctx, &keys, (size_t)nt));
        hipLaunchKernelGGL(k_u32_from_i32, dim3(tsg_div_up(nt, 256)), dim3(256), 0, ctx->stream, t->tile_colidx, nt, keys);
        hipLaunchKernelGGL(k_iota, dim3(tsg_div_up(nt, 256)), dim3(256), 0, ctx->stream, t->csc_tileid, nt);
        hipLaunchKernelGGL(k_count_keys, dim3(tsg_div_up(nt, 256)), dim3(256), 0, ctx->stream, t->tile_colidx, nt, t->csc_ptr);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u32(ctx, keys, t->csc_tileid, nt, bits_for(t->tilen)));
        hipLaunchKernelGGL(k_gather_i32, dim3(tsg_div_up(nt, 256)), dim3(256), 0, ctx->stream, t->tile_rowidx,
                           t->csc_tileid, nt, t->csc_rowidx);
        TSG_LAUNCH_CHECK();
        tsg_free(ctx, keys);
    }
    TSG_TRY(tsg_scan_i32(ctx, t->csc_ptr, t->csc_ptr, t->tilen, nullptr));
    return TSG_OK;
}

// ---------------------------------------------------------------- tile -> CSR
// Follows tile2csr (src/tile2csr.h:71-137): per-row counts from the tiles of
// the row's tile-row, exclusive scan, then entries copied tile by tile in
// ascending tile column (so columns come out sorted); explicit zeros kept.
// Row r of tile t spans [ptr[r], end) with end = tile nnz for the last row
// of the tile-row (rowlen-1) and ptr[r+1] otherwise (src/tile2csr.h:19-21),
// so only the offsets are needed, not the masks.
__device__ __forceinline__ int t2c_row_len(const int *tile_nnz, const unsigned char *csr_ptr, int t, int r,
                                           int rowlen) {
    const int s = csr_ptr[(long long)t * 16 + r];
    const int e = (r == rowlen - 1) ? tile_nnz[t + 1] - tile_nnz[t] : csr_ptr[(long long)t * 16 + r + 1];
    return e - s;
}

template <int PASS>
__global__ void __launch_bounds__(256)
k_t2c_rows(int m, const int *__restrict__ tile_ptr, const int *__restrict__ tile_colidx,
           const int *__restrict__ tile_nnz, const unsigned char *__restrict__ csr_ptr,
           const unsigned char *__restrict__ tcol, const double *__restrict__ tval,
           int *__restrict__ rowcnt_or_ptr, int *__restrict__ col, double *__restrict__ val, int jbase) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int I = i >> 4, r = i & 15;
    const int rowlen = min(16, m - I * 16);
    const int t0 = tile_ptr[I], t1 = tile_ptr[I + 1];
    if (PASS == 0) {
        int c = 0;
        for (int t = t0; t < t1; t++) c += t2c_row_len(tile_nnz, csr_ptr, t, r, rowlen);
        rowcnt_or_ptr[i] = c;
    } else {
        int out = rowcnt_or_ptr[i];
        for (int t = t0; t < t1; t++) {
            const int k = t2c_row_len(tile_nnz, csr_ptr, t, r, rowlen);
            if (k <= 0) continue;
            const int s = tile_nnz[t] + csr_ptr[(long long)t * 16 + r];
            const int J16 = (tile_colidx[t] - jbase) * 16;
            for (int j = 0; j < k; j++) {
                col[out + j] = J16 + tcol[s + j];
                val[out + j] = tval[s + j];
            }
            out += k;
        }
    }
}

extern "C" int tsg_tile2csr(tsg_ctx *ctx, const tsg_dtiles *t, tsg_dcsr *out) {
    return tsg_tile2csr_view(ctx, t, 0, out);
}

// t may be a view of a tile-row range of a larger tile matrix (tile_ptr
// pointing into the parent's, every other array the parent's, indices
// absolute); output columns are taken relative to tile column jbase.
int tsg_tile2csr_view(tsg_ctx *ctx, const tsg_dtiles *t, int jbase, tsg_dcsr *out) {
    memset(out, 0, sizeof(*out));
    out->m = t->m;
    out->n = t->n;
    TSG_TRY(tsg_alloc_t(ctx, &out->rowptr, (size_t)t->m + 1));
    if (t->m > 0) {
        hipLaunchKernelGGL(k_t2c_rows<0>, dim3(tsg_div_up(t->m, 256)), dim3(256), 0, ctx->stream, t->m, t->tile_ptr,
                           t->tile_colidx, t->tile_nnz, t->csr_ptr, t->csr_col, t->val, out->rowptr,
                           (int *)nullptr, (double *)nullptr, jbase);
        TSG_LAUNCH_CHECK();
    }
    long long nnz = 0;
    TSG_TRY(tsg_scan_i32(ctx, out->rowptr, out->rowptr, t->m, &nnz));
    out->nnz = (int)nnz;
    TSG_TRY(tsg_alloc_t(ctx, &out->colidx, (size_t)nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &out->val, (size_t)nnz + 1));
    if (t->m > 0 && nnz > 0) {
        hipLaunchKernelGGL(k_t2c_rows<1>, dim3(tsg_div_up(t->m, 256)), dim3(256), 0, ctx->stream, t->m, t->tile_ptr,
                           t->tile_colidx, t->tile_nnz, t->csr_ptr, t->csr_col, t->val, out->rowptr,
                           out->colidx, out->val, jbase);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

// ------------------------------------------------------------------ checksums, sampled rows
// splitmix64's finalizer: a bijective 64-bit mix.
__device__ __forceinline__ unsigned long long csum_mix(unsigned long long x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// One thread per row of the view: the row's entries (as tile2csr emits
// them) and its share of every tile of the tile-row (row mask and row
// offset; row 0 adds the tile's nonzero count).  Sums are order-free, so
// they add up over chunks.  acc[0] tiles, acc[1] entries, acc[2] nonzeros,
// acc[3] tiles counted, acc[4] entries with row and column swapped (equal
// to acc[1] exactly when C is bit-symmetric).
__global__ void __launch_bounds__(256)
k_checksum(int m, const int *__restrict__ tile_ptr, const int *__restrict__ tile_colidx,
           const int *__restrict__ tile_nnz, const unsigned char *__restrict__ csr_ptr,
           const unsigned short *__restrict__ mask, const unsigned char *__restrict__ tcol,
           const double *__restrict__ tval, long long row_base, int jbase, unsigned long long *__restrict__ acc) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    unsigned long long ht = 0, he = 0, hx = 0, nz = 0, nt = 0;
    if (i < m) {
        const int I = i >> 4, r = i & 15;
        const int rowlen = min(16, m - I * 16);
        const long long row = row_base + i, Ig = (row_base >> 4) + I;
        for (int t = tile_ptr[I]; t < tile_ptr[I + 1]; t++) {
            const long long J = tile_colidx[t] - jbase;
            const unsigned long long tk = csum_mix(((unsigned long long)Ig << 32) ^ (unsigned long long)J);
            ht += csum_mix(tk ^ ((unsigned long long)(r + 1) << 40) ^
                           ((unsigned long long)mask[(long long)t * 16 + r] << 8) ^ csr_ptr[(long long)t * 16 + r]);
            if (r == 0) {
                ht += csum_mix(tk + (unsigned long long)(tile_nnz[t + 1] - tile_nnz[t]));
                nt++;
            }
            const int k = t2c_row_len(tile_nnz, csr_ptr, t, r, rowlen);
            const int s = tile_nnz[t] + csr_ptr[(long long)t * 16 + r];
            for (int j = 0; j < k; j++) {
                const unsigned long long col = (unsigned long long)(J * 16 + tcol[s + j]);
                const unsigned long long vb = (unsigned long long)__double_as_longlong(tval[s + j]);
                he += csum_mix(csum_mix(((unsigned long long)row << 32) ^ col) ^ vb);
                hx += csum_mix(csum_mix((col << 32) ^ (unsigned long long)row) ^ vb);
            }
            nz += k > 0 ? k : 0;
        }
    }
    ht = tsg_wave_sum(ht);
    he = tsg_wave_sum(he);
    hx = tsg_wave_sum(hx);
    nz = tsg_wave_sum(nz);
    nt = tsg_wave_sum(nt);
    if (__lane_id() == 0) {
        atomicAdd(&acc[0], ht);
        atomicAdd(&acc[1], he);
        atomicAdd(&acc[2], nz);
        atomicAdd(&acc[3], nt);
        atomicAdd(&acc[4], hx);
    }
}

extern "C" int tsg_tiles_checksum(tsg_ctx *ctx, const tsg_dtiles *c, long long row_base, int jbase,
                                  tsg_csum *out) {
    memset(out, 0, sizeof(*out));
    if (row_base < 0 || (row_base & 15)) {
        tsg_set_error("tsg_tiles_checksum: row_base must be a non-negative multiple of 16");
        return TSG_ERR_ARG;
    }
    if (c->m <= 0) return TSG_OK;
    unsigned long long *acc;
    TSG_TRY(tsg_alloc_t(ctx, &acc, 5));
    int st = TSG_OK;
    if (hipMemsetAsync(acc, 0, 40, ctx->stream) != hipSuccess) {
        tsg_set_error("tsg_tiles_checksum: memset failed");
        st = TSG_ERR_HIP;
    }
    if (st == TSG_OK) {
        hipLaunchKernelGGL(k_checksum, dim3(tsg_div_up(c->m, 256)), dim3(256), 0, ctx->stream, c->m, c->tile_ptr,
                           c->tile_colidx, c->tile_nnz, c->csr_ptr, c->mask, c->csr_col, c->val, row_base, jbase,
                           acc);
        if (hipGetLastError() != hipSuccess) {
            tsg_set_error("tsg_tiles_checksum: launch failed");
            st = TSG_ERR_HIP;
        }
    }
    unsigned long long h[5] = {0, 0, 0, 0, 0};
    if (st == TSG_OK) st = tsg_d2h(ctx, h, acc, 40);
    if (st == TSG_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
        tsg_set_error("tsg_tiles_checksum: synchronize failed");
        st = TSG_ERR_HIP;
    }
    tsg_free(ctx, acc);
    TSG_TRY(st);
    out->tiles = h[0];
    out->entries = h[1];
    out->nnz = (long long)h[2];
    out->numtile = (long long)h[3];
    out->entries_t = h[4];
    return TSG_OK;
}

// Selected rows of a view (row ids local to the view, any order) as CSR.
template <int PASS>
__global__ void __launch_bounds__(256)
k_rows_sel(int m, int nsel, const int *__restrict__ sel, const int *__restrict__ tile_ptr,
           const int *__restrict__ tile_colidx, const int *__restrict__ tile_nnz,
           const unsigned char *__restrict__ csr_ptr, const unsigned char *__restrict__ tcol,
           const double *__restrict__ tval, int *__restrict__ cnt_or_ptr, int *__restrict__ col,
           double *__restrict__ val, int jbase) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= nsel) return;
    const int i = sel[x];
    const int I = i >> 4, r = i & 15;
    const int rowlen = min(16, m - I * 16);
    const int t0 = tile_ptr[I], t1 = tile_ptr[I + 1];
    if (PASS == 0) {
        int c = 0;
        for (int t = t0; t < t1; t++) c += t2c_row_len(tile_nnz, csr_ptr, t, r, rowlen);
        cnt_or_ptr[x] = c;
    } else {
        int out = cnt_or_ptr[x];
        for (int t = t0; t < t1; t++) {
            const int k = t2c_row_len(tile_nnz, csr_ptr, t, r, rowlen);
            if (k <= 0) continue;
            const int s = tile_nnz[t] + csr_ptr[(long long)t * 16 + r];
            const int J16 = (tile_colidx[t] - jbase) * 16;
            for (int j = 0; j < k; j++) {
                col[out + j] = J16 + tcol[s + j];
                val[out + j] = tval[s + j];
            }
            out += k;
        }
    }
}

extern "C" int tsg_tiles_rows(tsg_ctx *ctx, const tsg_dtiles *c, int jbase, const int *rows, int nrows,
                              tsg_hcsr *out) {
    memset(out, 0, sizeof(*out));
    if (nrows < 0 || (nrows > 0 && !rows)) { tsg_set_error("tsg_tiles_rows: bad row list"); return TSG_ERR_ARG; }
    for (int x = 0; x < nrows; x++)
        if (rows[x] < 0 || rows[x] >= c->m) {
            tsg_set_error("tsg_tiles_rows: row %d outside [0, %d)", rows[x], c->m);
            return TSG_ERR_ARG;
        }
    out->m = nrows;
    out->n = c->n;
    out->rowptr = (int *)malloc(4 * ((size_t)nrows + 1));
    if (!out->rowptr) { tsg_set_error("tsg_tiles_rows: host allocation failed"); return TSG_ERR_NOMEM; }
    out->rowptr[0] = 0;
    if (nrows == 0) {
        out->colidx = (int *)malloc(4);
        out->val = (double *)malloc(8);
        return TSG_OK;
    }
    int *dsel = nullptr, *dptr = nullptr, *dcol = nullptr;
    double *dval = nullptr;
    long long nnz = 0;
    const dim3 grid(tsg_div_up(nrows, 256));
    int st = tsg_alloc_t(ctx, &dsel, (size_t)nrows);
    if (st == TSG_OK) st = tsg_alloc_t(ctx, &dptr, (size_t)nrows + 1);
    if (st == TSG_OK) st = tsg_h2d(ctx, dsel, rows, 4 * (size_t)nrows);
    if (st == TSG_OK) {
        hipLaunchKernelGGL(k_rows_sel<0>, grid, dim3(256), 0, ctx->stream, c->m, nrows, dsel, c->tile_ptr,
                           c->tile_colidx, c->tile_nnz, c->csr_ptr, c->csr_col, c->val, dptr, (int *)nullptr,
                           (double *)nullptr, jbase);
        if (hipGetLastError() != hipSuccess) { tsg_set_error("tsg_tiles_rows: launch failed"); st = TSG_ERR_HIP; }
    }
    if (st == TSG_OK) st = tsg_scan_i32(ctx, dptr, dptr, nrows, &nnz);
    if (st == TSG_OK) st = tsg_alloc_t(ctx, &dcol, (size_t)nnz + 1);
    if (st == TSG_OK) st = tsg_alloc_t(ctx, &dval, (size_t)nnz + 1);
    if (st == TSG_OK && nnz > 0) {
        hipLaunchKernelGGL(k_rows_sel<1>, grid, dim3(256), 0, ctx->stream, c->m, nrows, dsel, c->tile_ptr,
                           c->tile_colidx, c->tile_nnz, c->csr_ptr, c->csr_col, c->val, dptr, dcol, dval, jbase);
        if (hipGetLastError() != hipSuccess) { tsg_set_error("tsg_tiles_rows: launch failed"); st = TSG_ERR_HIP; }
    }
    if (st == TSG_OK) {
        out->nnz = (int)nnz;
        out->colidx = (int *)malloc(4 * (size_t)(nnz ? nnz : 1));
        out->val = (double *)malloc(8 * (size_t)(nnz ? nnz : 1));
        if (!out->colidx || !out->val) { tsg_set_error("tsg_tiles_rows: host allocation failed"); st = TSG_ERR_NOMEM; }
    }
    if (st == TSG_OK) st = tsg_d2h(ctx, out->rowptr, dptr, 4 * ((size_t)nrows + 1));
    if (st == TSG_OK && nnz > 0) st = tsg_d2h(ctx, out->colidx, dcol, 4 * (size_t)nnz);
    if (st == TSG_OK && nnz > 0) st = tsg_d2h(ctx, out->val, dval, 8 * (size_t)nnz);
    if (st == TSG_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
        tsg_set_error("tsg_tiles_rows: synchronize failed");
        st = TSG_ERR_HIP;
    }
    tsg_free(ctx, dval);
    tsg_free(ctx, dcol);
    tsg_free(ctx, dptr);
    tsg_free(ctx, dsel);
    if (st != TSG_OK) tsg_hcsr_free(out);
    return st;
}

extern "C" int tsg_tiles_build_rowindex(tsg_ctx *ctx, tsg_dtiles *t) {
    if (t->rowx_ptr) return TSG_OK;
    tsg_dcsr d;
    TSG_TRY(tsg_tile2csr_view(ctx, t, 0, &d));
    t->rowx_ptr = d.rowptr;
    t->rowx_col = d.colidx;
    t->rowx_val = d.val;
    return TSG_OK;
}

extern "C" void tsg_tiles_free(tsg_ctx *ctx, tsg_dtiles *t) {
    if (!t) return;
    void *ps[] = {t->tile_ptr, t->tile_colidx, t->tile_rowidx, t->tile_nnz, t->csr_ptr, t->csr_col,
                  t->val, t->mask, t->csc_ptr, t->csc_rowidx, t->csc_tileid, t->rowx_ptr, t->rowx_col,
                  t->rowx_val};
    for (void *p : ps) tsg_free(ctx, p);
    memset(t, 0, sizeof(*t));
}

template <typename T>
static int dl(tsg_ctx *ctx, T **dst, const T *src, size_t n) {
    *dst = nullptr;
    if (!src) return TSG_OK;
    *dst = (T *)malloc((n ? n : 1) * sizeof(T));
    if (!*dst) { tsg_set_error("host allocation failed"); return TSG_ERR_NOMEM; }
    if (n) TSG_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    return TSG_OK;
}

extern "C" int tsg_tiles_download(tsg_ctx *ctx, const tsg_dtiles *d, tsg_htiles *h) {
    memset(h, 0, sizeof(*h));
    h->m = d->m; h->n = d->n; h->tilem = d->tilem; h->tilen = d->tilen;
    h->numtile = d->numtile; h->nnz = d->nnz;
    size_t nt = (size_t)d->numtile;
    TSG_TRY(dl(ctx, &h->tile_ptr, d->tile_ptr, (size_t)d->tilem + 1));
    TSG_TRY(dl(ctx, &h->tile_colidx, d->tile_colidx, nt));
    TSG_TRY(dl(ctx, &h->tile_rowidx, d->tile_rowidx, nt));
    TSG_TRY(dl(ctx, &h->tile_nnz, d->tile_nnz, nt + 1));
    TSG_TRY(dl(ctx, &h->csr_ptr, d->csr_ptr, nt * 16));
    TSG_TRY(dl(ctx, &h->csr_col, d->csr_col, (size_t)d->nnz));
    TSG_TRY(dl(ctx, &h->val, d->val, (size_t)d->nnz));
    TSG_TRY(dl(ctx, &h->mask, d->mask, nt * 16));
    TSG_TRY(dl(ctx, &h->csc_ptr, d->csc_ptr, (size_t)d->tilen + 1));
    TSG_TRY(dl(ctx, &h->csc_rowidx, d->csc_rowidx, nt));
    TSG_TRY(dl(ctx, &h->csc_tileid, d->csc_tileid, nt));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}

extern "C" void tsg_htiles_free(tsg_htiles *h) {
    if (!h) return;
    void *ps[] = {h->tile_ptr, h->tile_colidx, h->tile_rowidx, h->tile_nnz, h->csr_ptr, h->csr_col,
                  h->val, h->mask, h->csc_ptr, h->csc_rowidx, h->csc_tileid};
    for (void *p : ps) free(p);
    memset(h, 0, sizeof(*h));
}

// ---------------------------------------------------------------- transpose
__global__ void k_row_of_entry(const int *__restrict__ rowptr, int m, int *__restrict__ rowid) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    for (int e = rowptr[i]; e < rowptr[i + 1]; e++) rowid[e] = i;
}

// Stable CSR -> CSC (src/utils.h:158-195): a stable sort of the entries by
// column keeps rows ascending inside every column.
extern "C" int tsg_dcsr_transpose(tsg_ctx *ctx, const tsg_dcsr *a, tsg_dcsr *at) {
    memset(at, 0, sizeof(*at));
    TSG_TRY(tsg_validate_csr(ctx, a));
    at->m = a->n;
    at->n = a->m;
    at->nnz = a->nnz;
    const long long nnz = a->nnz;
    TSG_TRY(tsg_alloc_t(ctx, &at->rowptr, (size_t)a->n + 1));
    TSG_TRY(tsg_alloc_t(ctx, &at->colidx, (size_t)nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &at->val, (size_t)nnz + 1));
    TSG_HIP(hipMemsetAsync(at->rowptr, 0, ((size_t)a->n + 1) * 4, ctx->stream));
    if (nnz > 0) {
        unsigned *keys;
        int *perm, *rowid;
        TSG_TRY(tsg_alloc_t(ctx, &keys, (size_t)nnz));
        TSG_TRY(tsg_alloc_t(ctx, &perm, (size_t)nnz));
        TSG_TRY(tsg_alloc_t(ctx, &rowid, (size_t)nnz));
        const unsigned g = tsg_div_up(nnz, 256);
        hipLaunchKernelGGL(k_u32_from_i32, dim3(g), dim3(256), 0, ctx->stream, a->colidx, nnz, keys);
        hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, ctx->stream, perm, nnz);
        hipLaunchKernelGGL(k_count_keys, dim3(g), dim3(256), 0, ctx->stream, a->colidx, nnz, at->rowptr);
        hipLaunchKernelGGL(k_row_of_entry, dim3(tsg_div_up(a->m, 256)), dim3(256), 0, ctx->stream, a->rowptr, a->m, rowid);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_sort_pairs_u32(ctx, keys, perm, nnz, bits_for(a->n)));
        hipLaunchKernelGGL(k_gather_i32, dim3(g), dim3(256), 0, ctx->stream, rowid, perm, nnz, at->colidx);
        hipLaunchKernelGGL(k_gather_f64, dim3(g), dim3(256), 0, ctx->stream, a->val, perm, nnz, at->val);
        TSG_LAUNCH_CHECK();
        tsg_free(ctx, keys);
        tsg_free(ctx, perm);
        tsg_free(ctx, rowid);
    }
    TSG_TRY(tsg_scan_i32(ctx, at->rowptr, at->rowptr, a->n, nullptr));
    return TSG_OK;
}

// ---------------------------------------------------------------- nnzCub
__global__ void __launch_bounds__(256)
k_nnzcub(const int *__restrict__ colA, long long nnzA, const int *__restrict__ rowptrB,
         unsigned long long *__restrict__ out) {
    unsigned long long s = 0;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nnzA; e += (long long)gridDim.x * 256) {
        const int k = colA[e];
        s += (unsigned long long)(rowptrB[k + 1] - rowptrB[k]);
    }
    s = tsg_wave_sum(s);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

extern "C" int tsg_dcsr_nnzcub(tsg_ctx *ctx, const tsg_dcsr *a, const tsg_dcsr *b, unsigned long long *out) {
    if (a->n != b->m) { tsg_set_error("nnzcub: inner dimensions differ (%d vs %d)", a->n, b->m); return TSG_ERR_ARG; }
    TSG_TRY(tsg_validate_csr(ctx, a));
    TSG_TRY(tsg_validate_csr(ctx, b));
    void *scr;
    TSG_TRY(tsg_scratch(ctx, 64, &scr));
    TSG_HIP(hipMemsetAsync(scr, 0, 8, ctx->stream));
    if (a->nnz > 0) {
        unsigned nb = (unsigned)std::min<long long>(tsg_div_up(a->nnz, 256), 8192);
        hipLaunchKernelGGL(k_nnzcub, dim3(nb), dim3(256), 0, ctx->stream, a->colidx, (long long)a->nnz, b->rowptr,
                           (unsigned long long *)scr);
        TSG_LAUNCH_CHECK();
    }
    long long t;
    TSG_TRY(tsg_read_scalar_i64(ctx, scr, 1, &t));
    *out = (unsigned long long)t;
    return TSG_OK;
}

// ---------------------------------------------------------------- host-layer utilities
extern "C" int tsg_dmalloc(tsg_ctx *ctx, void **p, size_t bytes) { return tsg_alloc(ctx, p, bytes); }
extern "C" void tsg_dfree(tsg_ctx *ctx, void *p) { tsg_free(ctx, p); }
extern "C" int tsg_h2d(tsg_ctx *ctx, void *d, const void *h, size_t bytes) {
    if (bytes) TSG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}
extern "C" int tsg_d2h(tsg_ctx *ctx, void *h, const void *d, size_t bytes) {
    if (bytes) TSG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    return TSG_OK;
}
extern "C" int tsg_d_iota(tsg_ctx *ctx, int *d, long long n) {
    if (n > 0) {
        hipLaunchKernelGGL(k_iota, dim3(tsg_div_up(n, 256)), dim3(256), 0, ctx->stream, d, n);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

// (row << 4) | col in-tile bytes, the A-operand column layout of
// csr2tile_row_major (src/csr2tile.h:185).  One thread per (tile, row).
__global__ void __launch_bounds__(256)
k_pack_cols(long long nrows16, const int *__restrict__ tile_nnz, const unsigned char *__restrict__ ptr16,
            const unsigned char *__restrict__ col, unsigned char *__restrict__ out) {
    const long long x = (long long)blockIdx.x * 256 + threadIdx.x;
    if (x >= nrows16) return;
    const long long t = x >> 4;
    const int r = (int)(x & 15);
    const int tn = tile_nnz[t + 1] - tile_nnz[t];
    const int s = ptr16[x], e = r == 15 ? tn : ptr16[x + 1];
    for (int j = s; j < e; j++) out[tile_nnz[t] + j] = (unsigned char)((r << 4) | col[tile_nnz[t] + j]);
}

extern "C" int tsg_tiles_packed_cols(tsg_ctx *ctx, const tsg_dtiles *t, unsigned char *d_out) {
    const long long n16 = (long long)t->numtile * 16;
    if (n16 > 0) {
        hipLaunchKernelGGL(k_pack_cols, dim3(tsg_div_up(n16, 256)), dim3(256), 0, ctx->stream, n16, t->tile_nnz,
                           t->csr_ptr, t->csr_col, d_out);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

// Tile data re-ordered into CSC tile order (the B-operand layout of
// csr2tile_col_major, src/csr2tile.h:327-477).
__global__ void __launch_bounds__(256)
k_csc_meta(int nt, const int *__restrict__ cscid, const int *__restrict__ tile_nnz, const unsigned char *__restrict__ ptr16,
           const unsigned short *__restrict__ mask, unsigned char *__restrict__ optr16, unsigned short *__restrict__ omask,
           int *__restrict__ ocnt) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nt) return;
    const int s = cscid[p];
    *reinterpret_cast<uint4 *>(optr16 + (long long)p * 16) = *reinterpret_cast<const uint4 *>(ptr16 + (long long)s * 16);
    const uint4 *ms = reinterpret_cast<const uint4 *>(mask + (long long)s * 16);
    uint4 *md = reinterpret_cast<uint4 *>(omask + (long long)p * 16);
    md[0] = ms[0];
    md[1] = ms[1];
    ocnt[p] = tile_nnz[s + 1] - tile_nnz[s];
}
__global__ void __launch_bounds__(256)
k_csc_data(int nt, const int *__restrict__ cscid, const int *__restrict__ tile_nnz, const int *__restrict__ onnz,
           const unsigned char *__restrict__ col, const double *__restrict__ val, unsigned char *__restrict__ ocol,
           double *__restrict__ oval) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nt) return;
    const int s = cscid[p];
    const int src = tile_nnz[s], dst = onnz[p], n = onnz[p + 1] - dst;
    for (int j = 0; j < n; j++) {
        ocol[dst + j] = col[src + j];
        oval[dst + j] = val[src + j];
    }
}

extern "C" int tsg_tiles_csc_order(tsg_ctx *ctx, tsg_dtiles *t, tsg_dtiles *o) {
    memset(o, 0, sizeof(*o));
    TSG_TRY(tsg_tiles_build_csc(ctx, t));
    *o = *t;
    const long long nt = t->numtile;
    o->tile_ptr = o->tile_colidx = o->tile_rowidx = o->tile_nnz = nullptr;
    o->csr_ptr = o->csr_col = nullptr;
    o->val = nullptr;
    o->mask = nullptr;
    o->csc_ptr = o->csc_rowidx = o->csc_tileid = nullptr;
    o->rowx_ptr = o->rowx_col = nullptr;
    o->rowx_val = nullptr;
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_ptr, (size_t)t->tilem + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_colidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_rowidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csc_ptr, (size_t)t->tilen + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csc_rowidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csc_tileid, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_nnz, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csr_ptr, (size_t)nt * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &o->mask, (size_t)nt * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &o->csr_col, (size_t)t->nnz + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->val, (size_t)t->nnz + 1));
    hipStream_t s = ctx->stream;
    TSG_HIP(hipMemcpyAsync(o->tile_ptr, t->tile_ptr, ((size_t)t->tilem + 1) * 4, hipMemcpyDeviceToDevice, s));
    TSG_HIP(hipMemcpyAsync(o->csc_ptr, t->csc_ptr, ((size_t)t->tilen + 1) * 4, hipMemcpyDeviceToDevice, s));
    if (nt > 0) {
        TSG_HIP(hipMemcpyAsync(o->tile_colidx, t->tile_colidx, nt * 4, hipMemcpyDeviceToDevice, s));
        TSG_HIP(hipMemcpyAsync(o->tile_rowidx, t->tile_rowidx, nt * 4, hipMemcpyDeviceToDevice, s));
        TSG_HIP(hipMemcpyAsync(o->csc_rowidx, t->csc_rowidx, nt * 4, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_iota, dim3(tsg_div_up(nt, 256)), dim3(256), 0, s, o->csc_tileid, nt);
        hipLaunchKernelGGL(k_csc_meta, dim3(tsg_div_up(nt, 256)), dim3(256), 0, s, (int)nt, t->csc_tileid, t->tile_nnz,
                           t->csr_ptr, t->mask, o->csr_ptr, o->mask, o->tile_nnz);
        TSG_LAUNCH_CHECK();
    }
    TSG_TRY(tsg_scan_i32(ctx, o->tile_nnz, o->tile_nnz, nt, nullptr));
    if (nt > 0) {
        hipLaunchKernelGGL(k_csc_data, dim3(tsg_div_up(nt, 256)), dim3(256), 0, s, (int)nt, t->csc_tileid, t->tile_nnz,
                           o->tile_nnz, t->csr_col, t->val, o->csr_col, o->val);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

// Inverse of tsg_tiles_csc_order: tile data given in CSC tile order (the B
// layout csr2tile_col_major hands to the reference's tilespgemm) is moved to
// row-major tile order, and csc_tileid maps each CSC position to its
// row-major tile.  One thread per tile column walks its CSC entries and
// binary-searches each tile in its tile-row's (sorted) column list.
__global__ void __launch_bounds__(256)
k_csc_to_rm(int tilen, const int *__restrict__ csc_ptr, const int *__restrict__ csc_row,
            const int *__restrict__ tile_ptr, const int *__restrict__ tile_col, const int *__restrict__ nnz_csc,
            int *__restrict__ rm, int *__restrict__ cnt_rm, int *__restrict__ bad) {
    const int J = blockIdx.x * 256 + threadIdx.x;
    if (J >= tilen) return;
    for (int p = csc_ptr[J]; p < csc_ptr[J + 1]; p++) {
        const int K = csc_row[p];
        const int s = tile_ptr[K], len = tile_ptr[K + 1] - s;
        int lo = 0, hi = len - 1, f = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1, v = tile_col[s + mid];
            if (v == J) { f = mid; break; }
            if (v < J) lo = mid + 1; else hi = mid - 1;
        }
        if (f < 0) { atomicOr(bad, 1); rm[p] = 0; continue; }
        rm[p] = s + f;
        cnt_rm[s + f] = nnz_csc[p + 1] - nnz_csc[p];
    }
}
__global__ void __launch_bounds__(256)
k_scatter_meta(int nt, const int *__restrict__ rm, const unsigned char *__restrict__ ptr16,
               const unsigned short *__restrict__ mask, unsigned char *__restrict__ optr16,
               unsigned short *__restrict__ omask) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nt) return;
    const long long d = rm[p];
    *reinterpret_cast<uint4 *>(optr16 + d * 16) = *reinterpret_cast<const uint4 *>(ptr16 + (long long)p * 16);
    const uint4 *ms = reinterpret_cast<const uint4 *>(mask + (long long)p * 16);
    uint4 *md = reinterpret_cast<uint4 *>(omask + d * 16);
    md[0] = ms[0];
    md[1] = ms[1];
}
__global__ void __launch_bounds__(256)
k_scatter_data(int nt, const int *__restrict__ rm, const int *__restrict__ nnz_csc, const int *__restrict__ nnz_rm,
               const unsigned char *__restrict__ col, const double *__restrict__ val, unsigned char *__restrict__ ocol,
               double *__restrict__ oval) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nt) return;
    const int src = nnz_csc[p], n = nnz_csc[p + 1] - src, dst = nnz_rm[rm[p]];
    for (int j = 0; j < n; j++) {
        if (ocol) ocol[dst + j] = col[src + j];
        oval[dst + j] = val[src + j];
    }
}

extern "C" int tsg_tiles_from_csc_order(tsg_ctx *ctx, const tsg_dtiles *in, tsg_dtiles *o) {
    memset(o, 0, sizeof(*o));
    if (!in->csc_ptr || !in->csc_rowidx || !in->tile_ptr || !in->tile_colidx) {
        tsg_set_error("tsg_tiles_from_csc_order: needs the row-major and the CSC tile index");
        return TSG_ERR_ARG;
    }
    const long long nt = in->numtile;
    o->m = in->m; o->n = in->n; o->tilem = in->tilem; o->tilen = in->tilen;
    o->numtile = in->numtile; o->nnz = in->nnz;
    hipStream_t s = ctx->stream;
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_ptr, (size_t)in->tilem + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_colidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_rowidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csc_ptr, (size_t)in->tilen + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csc_rowidx, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csc_tileid, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->tile_nnz, (size_t)nt + 1));
    TSG_TRY(tsg_alloc_t(ctx, &o->csr_ptr, (size_t)nt * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &o->mask, (size_t)nt * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &o->val, (size_t)in->nnz + 1));
    if (in->csr_col) TSG_TRY(tsg_alloc_t(ctx, &o->csr_col, (size_t)in->nnz + 1));
    TSG_HIP(hipMemcpyAsync(o->tile_ptr, in->tile_ptr, ((size_t)in->tilem + 1) * 4, hipMemcpyDeviceToDevice, s));
    TSG_HIP(hipMemcpyAsync(o->csc_ptr, in->csc_ptr, ((size_t)in->tilen + 1) * 4, hipMemcpyDeviceToDevice, s));
    void *scr;
    TSG_TRY(tsg_scratch(ctx, 64, &scr));
    TSG_HIP(hipMemsetAsync(scr, 0, 4, s));
    if (nt > 0) {
        TSG_HIP(hipMemcpyAsync(o->tile_colidx, in->tile_colidx, nt * 4, hipMemcpyDeviceToDevice, s));
        TSG_HIP(hipMemcpyAsync(o->csc_rowidx, in->csc_rowidx, nt * 4, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_csc_to_rm, dim3(tsg_div_up(in->tilen, 256)), dim3(256), 0, s, in->tilen, in->csc_ptr,
                           in->csc_rowidx, in->tile_ptr, in->tile_colidx, in->tile_nnz, o->csc_tileid, o->tile_nnz,
                           (int *)scr);
        TSG_LAUNCH_CHECK();
    }
    long long bad = 0;
    TSG_TRY(tsg_read_scalar_i64(ctx, scr, 0, &bad));
    if (bad) { tsg_set_error("CSC tile index does not match the row-major tile index"); return TSG_ERR_ARG; }
    TSG_TRY(tsg_scan_i32(ctx, o->tile_nnz, o->tile_nnz, nt, nullptr));
    if (nt > 0) {
        // tile row of every row-major tile
        hipLaunchKernelGGL(k_scatter_i32, dim3(tsg_div_up(nt, 256)), dim3(256), 0, s, in->csc_rowidx, o->csc_tileid, nt,
                           o->tile_rowidx);
        hipLaunchKernelGGL(k_scatter_meta, dim3(tsg_div_up(nt, 256)), dim3(256), 0, s, (int)nt, o->csc_tileid,
                           in->csr_ptr, in->mask, o->csr_ptr, o->mask);
        hipLaunchKernelGGL(k_scatter_data, dim3(tsg_div_up(nt, 256)), dim3(256), 0, s, (int)nt, o->csc_tileid,
                           in->tile_nnz, o->tile_nnz, in->csr_col, in->val, o->csr_col, o->val);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}
