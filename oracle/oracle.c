/*
 * oracle.c -- CPU oracle for the TileSpGEMM hot path.  TEST INFRASTRUCTURE
 * ONLY: the product library never links, loads or calls this file; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker / the timed CPU baseline.
 *
 * Each function restates one reference routine (paths relative to the
 * reference repository root) in plain C/OpenMP:
 *
 *   ora_spa_structure   src/spgemm_serialref_spa_new.h:7-105  (spgemm_spa:
 *                       per-row (nB/32+1)-word bitmap, count pass then fill
 *                       pass; structure only)
 *   ora_spa_values      src/external/cusparse/spgemm_serialref_spa.h:7-31
 *                       (compute_dense_row: acc[j] += a_ik * b_kj in A-row
 *                       order then B-row order; the contraction the
 *                       reference build applies with -mfma, src/Makefile:29,
 *                       is made explicit as fma(a, b, acc))
 *   ora_spgemm_rowlist  the same, for a list of rows (sampled-row checks of
 *                       products too large for the whole oracle)
 *   ora_nnzcub          src/main.cu:391-397
 *   ora_csr2tile        src/csr2tile.h:6-259 (row-major tiles; in-tile
 *                       columns plain or packed (row<<4)|col)
 *   ora_step1_tiles     the tile-level boolean product computed by
 *                       src/tilespgemm-cuda.h:257-369 (C tile list of every
 *                       tile-row, sorted; includes structurally empty tiles)
 *   ora_c_tile_masks    the per-tile row masks / u8 row offsets / tile nnz of
 *                       step 2 (src/tilespgemm-cuda.h:658-700), derived from
 *                       the element-level CSR of C on the step-1 tile list
 *   ora_transpose       src/utils.h:158-195 (stable counting-sort transpose)
 *   ora_degree_order    reorder_algorithm/common/order_utils.py:78-81
 *   ora_rcm             Cuthill-McKee as structured in the reference's
 *                       SciPy-based native_rcm (order_utils.py:84-104) with
 *                       every tie broken by (degree, vertex id); see header
 *   ora_permute_csr     src/reorder.h:2429-2551 (apply / restore)
 *
 * Pinning (tests/test_oracle.py): closed-form stencil counts, SciPy sparse
 * products (an independent SpGEMM: structure and integer-valued results
 * bit-exact, float values to 1e-12 of sum|a*b|), brute-force dense products
 * with exact rational fma on tiny inputs, scipy RCM on tie-free graphs and
 * RCM invariants.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ SPA */
static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return x < y ? -1 : x > y;
}

/* mode 1: count only (rowptrC gets the exclusive scan, *nnzC the total)
 * mode 0: fill colC using rowptrC from a previous count pass.
 * Faithful to the reference: a full (nB/32+1)-word bitmap per row, cleared
 * and scanned whole, so the cost is O(m * nB / 32) regardless of sparsity. */
long long ora_spa_structure(int mA, const int *rowptrA, const int *colA, const int *rowptrB,
                            const int *colB, int nB, long long *rowptrC, int *colC, int mode) {
    const int nwords = nB / 32 + 1;
    int nthreads = omp_get_max_threads();
    unsigned *flag_g = (unsigned *)malloc((size_t)nthreads * nwords * sizeof(unsigned));
    if (!flag_g) return -1;
    if (mode == 1) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < mA; i++) {
            unsigned *flag = flag_g + (size_t)omp_get_thread_num() * nwords;
            memset(flag, 0, sizeof(unsigned) * nwords);
            for (int p = rowptrA[i]; p < rowptrA[i + 1]; p++) {
                const int k = colA[p];
                for (int q = rowptrB[k]; q < rowptrB[k + 1]; q++) {
                    const int key = colB[q];
                    flag[key / 32] |= 1u << (key % 32);
                }
            }
            long long c = 0;
            for (int w = 0; w < nwords; w++) c += __builtin_popcount(flag[w]);
            rowptrC[i] = c;
        }
        long long s = 0;
        for (int i = 0; i < mA; i++) {
            long long v = rowptrC[i];
            rowptrC[i] = s;
            s += v;
        }
        rowptrC[mA] = s;
        free(flag_g);
        return s;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < mA; i++) {
        unsigned *flag = flag_g + (size_t)omp_get_thread_num() * nwords;
        memset(flag, 0, sizeof(unsigned) * nwords);
        for (int p = rowptrA[i]; p < rowptrA[i + 1]; p++) {
            const int k = colA[p];
            for (int q = rowptrB[k]; q < rowptrB[k + 1]; q++) {
                const int key = colB[q];
                flag[key / 32] |= 1u << (key % 32);
            }
        }
        long long out = rowptrC[i];
        for (int w = 0; w < nwords; w++) {
            unsigned x = flag[w];
            while (x) {
                int b = __builtin_ctz(x);
                x &= x - 1;
                colC[out++] = w * 32 + b;
            }
        }
    }
    free(flag_g);
    return rowptrC[mA];
}

/* Values on a given structure (sorted colC per row).  Dense accumulator of
 * nB doubles per thread, reset only where touched. */
int ora_spa_values(int mA, const int *rowptrA, const int *colA, const double *valA, const int *rowptrB,
                   const int *colB, const double *valB, int nB, const long long *rowptrC, const int *colC,
                   double *valC) {
    int nthreads = omp_get_max_threads();
    double *acc_g = (double *)calloc((size_t)nthreads * nB + 1, sizeof(double));
    if (!acc_g) return -1;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < mA; i++) {
        double *acc = acc_g + (size_t)omp_get_thread_num() * nB;
        for (int p = rowptrA[i]; p < rowptrA[i + 1]; p++) {
            const int k = colA[p];
            const double a = valA[p];
            for (int q = rowptrB[k]; q < rowptrB[k + 1]; q++) acc[colB[q]] = fma(a, valB[q], acc[colB[q]]);
        }
        for (long long e = rowptrC[i]; e < rowptrC[i + 1]; e++) {
            valC[e] = acc[colC[e]];
            acc[colC[e]] = 0.0;
        }
    }
    free(acc_g);
    return 0;
}

/* Structure + values of C restricted to rows [r0, r1) (sampled-row checks);
 * returns nnz of the slice, -1 on allocation failure.  rowptrC has
 * r1-r0+1 entries; colC/valC may be NULL for a count-only call. */
long long ora_spgemm_rows(int r0, int r1, const int *rowptrA, const int *colA, const double *valA,
                          const int *rowptrB, const int *colB, const double *valB, int nB, long long *rowptrC,
                          int *colC, double *valC) {
    unsigned char *seen = (unsigned char *)calloc((size_t)nB + 1, 1);
    double *acc = (double *)calloc((size_t)nB + 1, sizeof(double));
    int *list = (int *)malloc(((size_t)nB + 1) * sizeof(int));
    if (!seen || !acc || !list) { free(seen); free(acc); free(list); return -1; }
    long long out = 0;
    rowptrC[0] = 0;
    for (int i = r0; i < r1; i++) {
        int nl = 0;
        for (int p = rowptrA[i]; p < rowptrA[i + 1]; p++) {
            const int k = colA[p];
            const double a = valA ? valA[p] : 0.0;
            for (int q = rowptrB[k]; q < rowptrB[k + 1]; q++) {
                const int j = colB[q];
                if (!seen[j]) { seen[j] = 1; list[nl++] = j; }
                if (valA) acc[j] = fma(a, valB[q], acc[j]);
            }
        }
        qsort(list, nl, sizeof(int), cmp_int);
        for (int x = 0; x < nl; x++) {
            const int j = list[x];
            if (colC) colC[out + x] = j;
            if (valC) valC[out + x] = acc[j];
            acc[j] = 0.0;
            seen[j] = 0;
        }
        out += nl;
        rowptrC[i - r0 + 1] = out;
    }
    free(seen); free(acc); free(list);
    return out;
}

/* Rows rows[0..nrows) of C = A*B, the same per-row restatement as
 * ora_spgemm_rows (compute_dense_row order, src/external/cusparse/
 * spgemm_serialref_spa.h:18-29), the rows spread over OpenMP threads.
 * colC == NULL: per-row counts into cnt[x] (return = total).  Otherwise
 * rowptrC[x] (from a count pass) places row x's sorted columns/values. */
long long ora_spgemm_rowlist(int nrows, const int *rows, const int *rowptrA, const int *colA, const double *valA,
                             const int *rowptrB, const int *colB, const double *valB, int nB, long long *cnt,
                             const long long *rowptrC, int *colC, double *valC) {
    long long total = 0;
    int fail = 0;
#pragma omp parallel reduction(+ : total) reduction(| : fail)
    {
        unsigned char *seen = (unsigned char *)calloc((size_t)nB + 1, 1);
        double *acc = (double *)calloc((size_t)nB + 1, sizeof(double));
        int *list = (int *)malloc(((size_t)nB + 1) * sizeof(int));
        if (!seen || !acc || !list) {
            fail = 1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int x = 0; x < nrows; x++) {
                const int i = rows[x];
                int nl = 0;
                for (int p = rowptrA[i]; p < rowptrA[i + 1]; p++) {
                    const int k = colA[p];
                    const double a = valA ? valA[p] : 0.0;
                    for (int q = rowptrB[k]; q < rowptrB[k + 1]; q++) {
                        const int j = colB[q];
                        if (!seen[j]) { seen[j] = 1; list[nl++] = j; }
                        if (valA && colC) acc[j] = fma(a, valB[q], acc[j]);
                    }
                }
                if (colC) qsort(list, nl, sizeof(int), cmp_int);
                for (int y = 0; y < nl; y++) {
                    const int j = list[y];
                    if (colC) colC[rowptrC[x] + y] = j;
                    if (colC && valC) valC[rowptrC[x] + y] = acc[j];
                    acc[j] = 0.0;
                    seen[j] = 0;
                }
                if (!colC) cnt[x] = nl;
                total += nl;
            }
        }
        free(seen); free(acc); free(list);
    }
    return fail ? -1 : total;
}

unsigned long long ora_nnzcub(int nnzA, const int *colA, const int *rowptrB) {
    unsigned long long s = 0;
    for (int e = 0; e < nnzA; e++) s += (unsigned long long)(rowptrB[colA[e] + 1] - rowptrB[colA[e]]);
    return s;
}

/* ------------------------------------------------------------------ tiles */
/* Pass 1: numtile (return) and tile_ptr[tilem+1].  Pass 2 (tile_colidx !=
 * NULL): everything else.  packed: tile_csr_Col = (row<<4)|col as for A. */

long long ora_csr2tile(int m, int n, const int *rowptr, const int *col, const double *val, int packed,
                       int *tile_ptr, int *tile_colidx, int *tile_rowidx, int *tile_nnz, unsigned char *csr_ptr,
                       unsigned char *csr_col, double *tval, unsigned short *mask) {
    const int tilem = (m + 15) / 16, tilen = (n + 15) / 16;
    int *flag = (int *)malloc(((size_t)tilen + 1) * sizeof(int));
    int *list = (int *)malloc(((size_t)tilen + 1) * sizeof(int));
    if (!flag || !list) { free(flag); free(list); return -1; }
    for (int j = 0; j < tilen; j++) flag[j] = -1;
    long long nt = 0;
    tile_ptr[0] = 0;
    for (int I = 0; I < tilem; I++) {
        const int r0 = I * 16, r1 = (r0 + 16 < m) ? r0 + 16 : m;
        int nl = 0;
        for (int e = rowptr[r0]; e < rowptr[r1]; e++) {
            int J = col[e] / 16;
            if (flag[J] != I) { flag[J] = I; list[nl++] = J; }
        }
        qsort(list, nl, sizeof(int), cmp_int);
        if (tile_colidx) {
            /* per tile: counts per row, masks */
            for (int x = 0; x < nl; x++) {
                long long t = nt + x;
                tile_colidx[t] = list[x];
                tile_rowidx[t] = I;
                memset(csr_ptr + t * 16, 0, 16);
                memset(mask + t * 16, 0, 32);
            }
            for (int i = r0; i < r1; i++)
                for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
                    int J = col[e] / 16;
                    /* binary search J in list */
                    int lo = 0, hi = nl - 1;
                    while (lo < hi) { int mid = (lo + hi) / 2; if (list[mid] < J) lo = mid + 1; else hi = mid; }
                    long long t = nt + lo;
                    csr_ptr[t * 16 + (i - r0)]++;
                    mask[t * 16 + (i - r0)] |= (unsigned short)(0x8000u >> (col[e] % 16));
                }
            for (int x = 0; x < nl; x++) {
                long long t = nt + x;
                int s = 0;
                for (int r = 0; r < 16; r++) { int c = csr_ptr[t * 16 + r]; csr_ptr[t * 16 + r] = (unsigned char)s; s += c; }
                tile_nnz[t] = s;
            }
        }
        nt += nl;
        tile_ptr[I + 1] = (int)nt;
    }
    if (tile_colidx) {
        long long s = 0;
        for (long long t = 0; t < nt; t++) { long long v = tile_nnz[t]; tile_nnz[t] = (int)s; s += v; }
        tile_nnz[nt] = (int)s;
        /* scatter values: row-major inside tile, ascending column (CSR order) */
        for (int I = 0; I < tilem; I++) {
            const int r0 = I * 16, r1 = (r0 + 16 < m) ? r0 + 16 : m;
            const int t0 = tile_ptr[I], t1 = tile_ptr[I + 1];
            for (int i = r0; i < r1; i++) {
                int t = t0, k = 0, curJ = -1;
                for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
                    int J = col[e] / 16;
                    if (J != curJ) { while (t < t1 && tile_colidx[t] < J) t++; curJ = J; k = 0; }
                    long long pos = (long long)tile_nnz[t] + csr_ptr[(long long)t * 16 + (i - r0)] + k;
                    tval[pos] = val[e];
                    csr_col[pos] = (unsigned char)(packed ? (((i - r0) << 4) | (col[e] % 16)) : (col[e] % 16));
                    k++;
                }
            }
        }
    }
    free(flag);
    free(list);
    return nt;
}

/* Step-1 tile list: for each A tile-row I, the sorted union of B's tile-row
 * lists over the tile columns K of A's tile-row.  Pass 1 (Ccol NULL): C tile
 * pointer + count; pass 2: columns. */
long long ora_step1_tiles(int tilemA, const int *Aptr, const int *Acol, int tilenB, const int *Bptr,
                          const int *Bcol, int *Cptr, int *Ccol) {
    int *flag = (int *)malloc(((size_t)tilenB + 1) * sizeof(int));
    int *list = (int *)malloc(((size_t)tilenB + 1) * sizeof(int));
    if (!flag || !list) { free(flag); free(list); return -1; }
    for (int j = 0; j < tilenB; j++) flag[j] = -1;
    long long nt = 0;
    Cptr[0] = 0;
    for (int I = 0; I < tilemA; I++) {
        int nl = 0;
        for (int p = Aptr[I]; p < Aptr[I + 1]; p++) {
            int K = Acol[p];
            for (int q = Bptr[K]; q < Bptr[K + 1]; q++) {
                int J = Bcol[q];
                if (flag[J] != I) { flag[J] = I; list[nl++] = J; }
            }
        }
        if (Ccol) {
            qsort(list, nl, sizeof(int), cmp_int);
            memcpy(Ccol + nt, list, (size_t)nl * sizeof(int));
        }
        nt += nl;
        Cptr[I + 1] = (int)nt;
    }
    free(flag);
    free(list);
    return nt;
}

/* Step-2 quantities on the step-1 tile list, from C's element CSR:
 * mask/ptr per (tile, row), tile nnz exclusive offsets.  Tiles with no
 * element (step-1 phantoms) get zero masks, zero offsets and nnz 0. */
int ora_c_tile_masks(int m, int tilem, const int *Cptr, const int *Ccol, long long numtile, const long long *rowptrC,
                     const int *colC, unsigned short *mask, unsigned char *ptr16, long long *tile_nnz) {
    memset(mask, 0, (size_t)numtile * 32);
    memset(ptr16, 0, (size_t)numtile * 16);
    long long *cnt = (long long *)calloc((size_t)numtile + 1, sizeof(long long));
    if (!cnt) return -1;
    for (int i = 0; i < m; i++) {
        const int I = i / 16, r = i % 16;
        int t = Cptr[I];
        for (long long e = rowptrC[i]; e < rowptrC[i + 1]; e++) {
            int J = colC[e] / 16;
            while (t < Cptr[I + 1] && Ccol[t] < J) t++;
            if (t >= Cptr[I + 1] || Ccol[t] != J) { free(cnt); return -2; } /* element outside step-1 tiles */
            mask[(long long)t * 16 + r] |= (unsigned short)(0x8000u >> (colC[e] % 16));
        }
    }
    (void)tilem;
    for (long long t = 0; t < numtile; t++) {
        int s = 0;
        for (int r = 0; r < 16; r++) { ptr16[t * 16 + r] = (unsigned char)s; s += __builtin_popcount(mask[t * 16 + r]); }
        cnt[t] = s;
    }
    long long s = 0;
    for (long long t = 0; t < numtile; t++) { tile_nnz[t] = s; s += cnt[t]; }
    tile_nnz[numtile] = s;
    free(cnt);
    return 0;
}

/* ------------------------------------------------------------------ transpose */
int ora_transpose(int m, int n, const int *rowptr, const int *col, const double *val, int *tptr, int *tcol,
                  double *tval) {
    memset(tptr, 0, sizeof(int) * ((size_t)n + 1));
    const int nnz = rowptr[m];
    for (int e = 0; e < nnz; e++) tptr[col[e] + 1]++;
    for (int j = 0; j < n; j++) tptr[j + 1] += tptr[j];
    int *cur = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    if (!cur) return -1;
    memcpy(cur, tptr, sizeof(int) * ((size_t)n + 1));
    for (int i = 0; i < m; i++)
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            int p = cur[col[e]]++;
            tcol[p] = i;
            if (val) tval[p] = val[e];
        }
    free(cur);
    return 0;
}

/* ------------------------------------------------------------------ orderings */
/* degree order: ids sorted by (-deg, id); deg = row length, plus the
 * mirrored count when `symmetric_header` (order_utils.py:41-45,78-81). */
typedef struct { long long k1; int id; } okey;
static int cmp_okey(const void *a, const void *b) {
    const okey *x = (const okey *)a, *y = (const okey *)b;
    if (x->k1 != y->k1) return x->k1 < y->k1 ? -1 : 1;
    return x->id < y->id ? -1 : x->id > y->id;
}

int ora_degree_order(int n, const int *rowptr, const int *col, int symmetric_header, int *old_id) {
    long long *deg = (long long *)calloc((size_t)n, sizeof(long long));
    okey *k = (okey *)malloc(sizeof(okey) * (size_t)(n ? n : 1));
    if (!deg || !k) { free(deg); free(k); return -1; }
    for (int i = 0; i < n; i++)
        for (int e = rowptr[i]; e < rowptr[i + 1]; e++) {
            deg[i]++;
            if (symmetric_header && col[e] != i) deg[col[e]]++;
        }
    for (int i = 0; i < n; i++) { k[i].k1 = -deg[i]; k[i].id = i; }
    qsort(k, n, sizeof(okey), cmp_okey);
    for (int i = 0; i < n; i++) old_id[i] = k[i].id;
    free(deg);
    free(k);
    return 0;
}

/* RCM on the symmetrised pattern G = pattern(A + A^T) (self loops kept, as
 * the reference builds its adjacency sets).  Degree as SciPy's
 * _node_degrees: row length of G plus one when the diagonal is present.
 * Cuthill-McKee: seeds taken in ascending (degree, id); BFS appends each
 * parent's unvisited neighbours sorted by (degree, id); the final order is
 * reversed.  gptr/gcol: the symmetrised graph (sorted rows) built by the
 * caller with ora_symmetrize. */
int ora_symmetrize(int n, const int *rowptr, const int *col, int *gptr, int *gcol, int cap) {
    /* count */
    int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
    int *tptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int nnz = rowptr[n];
    int *tcol = (int *)malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    if (!cnt || !tptr || !tcol) { free(cnt); free(tptr); free(tcol); return -1; }
    ora_transpose(n, n, rowptr, col, NULL, tptr, tcol, NULL);
    /* merge row i of A and row i of A^T (both sorted) */
    int out = 0;
    gptr[0] = 0;
    for (int i = 0; i < n; i++) {
        int a = rowptr[i], ae = rowptr[i + 1], b = tptr[i], be = tptr[i + 1];
        while (a < ae || b < be) {
            int v;
            if (b >= be || (a < ae && col[a] < tcol[b])) v = col[a++];
            else if (a >= ae || tcol[b] < col[a]) v = tcol[b++];
            else { v = col[a]; a++; b++; }
            if (out >= cap) { free(cnt); free(tptr); free(tcol); return -2; }
            gcol[out++] = v;
        }
        gptr[i + 1] = out;
    }
    free(cnt); free(tptr); free(tcol);
    return out;
}

int ora_rcm(int n, const int *gptr, const int *gcol, int *old_id) {
    long long *deg = (long long *)malloc(sizeof(long long) * (size_t)(n ? n : 1));
    okey *k = (okey *)malloc(sizeof(okey) * (size_t)(n ? n : 1));
    unsigned char *vis = (unsigned char *)calloc((size_t)n + 1, 1);
    int *order = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
    okey *kids = (okey *)malloc(sizeof(okey) * (size_t)(n ? n : 1));
    if (!deg || !k || !vis || !order || !kids) { free(deg); free(k); free(vis); free(order); free(kids); return -1; }
    for (int i = 0; i < n; i++) {
        deg[i] = gptr[i + 1] - gptr[i];
        for (int e = gptr[i]; e < gptr[i + 1]; e++) if (gcol[e] == i) { deg[i]++; break; }
        k[i].k1 = deg[i];
        k[i].id = i;
    }
    qsort(k, n, sizeof(okey), cmp_okey);
    int N = 0;
    for (int z = 0; z < n; z++) {
        int seed = k[z].id;
        if (vis[seed]) continue;
        vis[seed] = 1;
        order[N++] = seed;
        int head = N - 1;
        while (head < N) {
            int p = order[head++];
            int nk = 0;
            for (int e = gptr[p]; e < gptr[p + 1]; e++) {
                int v = gcol[e];
                if (!vis[v]) { vis[v] = 1; kids[nk].k1 = deg[v]; kids[nk].id = v; nk++; }
            }
            qsort(kids, nk, sizeof(okey), cmp_okey);
            for (int x = 0; x < nk; x++) order[N++] = kids[x].id;
        }
    }
    for (int i = 0; i < n; i++) old_id[i] = order[n - 1 - i];
    free(deg); free(k); free(vis); free(order); free(kids);
    return 0;
}

/* out = P * in * Q^T with out row r = in row row_old[r] and columns
 * relabelled by col_new[], rows sorted by column (src/reorder.h:2429-2504);
 * restore is the same call with the inverse maps (src/reorder.h:2506-2551). */
typedef struct { int c; double v; } cv;
static int cmp_cv(const void *a, const void *b) {
    int x = ((const cv *)a)->c, y = ((const cv *)b)->c;
    return x < y ? -1 : x > y;
}
int ora_permute_csr(int m, const int *rowptr, const int *col, const double *val, const int *row_old,
                    const int *col_new, int *optr, int *ocol, double *oval) {
    optr[0] = 0;
    for (int r = 0; r < m; r++) optr[r + 1] = optr[r] + (rowptr[row_old[r] + 1] - rowptr[row_old[r]]);
    int maxlen = 0;
    for (int r = 0; r < m; r++) if (optr[r + 1] - optr[r] > maxlen) maxlen = optr[r + 1] - optr[r];
    cv *buf = (cv *)malloc(sizeof(cv) * (size_t)(maxlen ? maxlen : 1));
    if (!buf) return -1;
    for (int r = 0; r < m; r++) {
        int o = row_old[r], L = 0;
        for (int e = rowptr[o]; e < rowptr[o + 1]; e++) { buf[L].c = col_new[col[e]]; buf[L].v = val ? val[e] : 0.0; L++; }
        qsort(buf, L, sizeof(cv), cmp_cv);
        for (int x = 0; x < L; x++) { ocol[optr[r] + x] = buf[x].c; if (oval) oval[optr[r] + x] = buf[x].v; }
    }
    free(buf);
    return 0;
}

int ora_num_threads(void) { return omp_get_max_threads(); }
