/*
 * compat.c -- the reference driver's entry points (include/tilespgemm.h)
 * implemented in C on top of the device C ABI (include/tsg.h).
 *
 * The reference computes csr2tile / tile2csr / transposes on the host with
 * OpenMP (src/csr2tile.h, src/tile2csr.h, src/utils.h) and only steps 1-3
 * on the GPU.  Here every function uploads its operands, runs the device
 * kernels and hands back host arrays in the reference layout, so the
 * driver's data flow and ownership rules (src/main.cu:135-638) are
 * unchanged.  A process-wide context is created on first use on the device
 * chosen by tsg_compat_set_device() (default: TSG_DEVICE or 0).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tilespgemm.h"
#include "../../include/tsg.h"

static tsg_ctx *g_ctx = NULL;
static int g_dev = -1;

int tsg_compat_set_device(int device) {
    if (g_ctx) {
        tsg_ctx_destroy(g_ctx);
        g_ctx = NULL;
    }
    g_dev = device;
    return 0;
}

/* The process-wide context, or NULL when no device can be opened. */
tsg_ctx *tsg_compat_ctx_try(void) {
    if (!g_ctx) {
        int dev = g_dev;
        if (dev < 0) {
            const char *e = getenv("TSG_DEVICE");
            dev = e ? atoi(e) : 0;
        }
        if (tsg_ctx_create(dev, &g_ctx) != TSG_OK) g_ctx = NULL;
    }
    return g_ctx;
}

tsg_ctx *tsg_compat_ctx(void) {
    tsg_ctx *c = tsg_compat_ctx_try();
    if (!c) {
        fprintf(stderr, "tilespgemm: cannot create device context: %s\n", tsg_last_error());
        abort();
    }
    return c;
}

/* Entry points returning void in the reference abort on failure. */
static void die(const char *where, int st) {
    fprintf(stderr, "tilespgemm: %s failed (%d): %s\n", where, st, tsg_last_error());
    abort();
}
#define CK(where, expr)                     \
    do {                                    \
        int st_ = (expr);                   \
        if (st_ != TSG_OK) die(where, st_); \
    } while (0)

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "tilespgemm: host allocation of %zu bytes failed\n", n);
        abort();
    }
    return p;
}

static void upload_csr(tsg_ctx *c, const SMatrix *m, tsg_dcsr *d) {
    tsg_hcsr h;
    h.m = m->m;
    h.n = m->n;
    h.nnz = m->nnz;
    h.rowptr = m->rowpointer;
    h.colidx = m->columnindex;
    h.val = m->value;
    CK("upload CSR", tsg_dcsr_upload(c, &h, d));
}

static void *fetch(tsg_ctx *c, const void *d, size_t bytes) {
    void *h = xmalloc(bytes);
    CK("download", tsg_d2h(c, h, d, bytes));
    return h;
}

void csr2tile_row_major(SMatrix *matrix) {
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dcsr d;
    tsg_dtiles t;
    upload_csr(c, matrix, &d);
    CK("csr2tile", tsg_csr2tile(c, &d, &t));
    unsigned char *packed = NULL;
    CK("alloc", tsg_dmalloc(c, (void **)&packed, (size_t)t.nnz + 1));
    CK("pack", tsg_tiles_packed_cols(c, &t, packed));
    const size_t nt = (size_t)t.numtile;
    matrix->tilem = t.tilem;
    matrix->tilen = t.tilen;
    matrix->numtile = t.numtile;
    matrix->tile_ptr = (int *)fetch(c, t.tile_ptr, ((size_t)t.tilem + 1) * 4);
    matrix->tile_columnidx = (int *)fetch(c, t.tile_colidx, nt * 4);
    matrix->tile_rowidx = (int *)fetch(c, t.tile_rowidx, nt * 4);
    matrix->tile_nnz = (int *)fetch(c, t.tile_nnz, (nt + 1) * 4);
    matrix->tile_csr_Ptr = (unsigned char *)fetch(c, t.csr_ptr, nt * 16);
    matrix->tile_csr_Col = (unsigned char *)fetch(c, packed, (size_t)t.nnz);
    matrix->tile_csr_Value = (double *)fetch(c, t.val, (size_t)t.nnz * 8);
    matrix->mask = (unsigned short *)fetch(c, t.mask, nt * 32);
    matrix->csc_tile_ptr = NULL;
    matrix->csc_tile_rowidx = NULL;
    tsg_dfree(c, packed);
    tsg_tiles_free(c, &t);
    tsg_dcsr_free(c, &d);
}

void csr2tile_col_major(SMatrix *matrix) {
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dcsr d;
    tsg_dtiles t, o;
    upload_csr(c, matrix, &d);
    CK("csr2tile", tsg_csr2tile(c, &d, &t));
    CK("csc order", tsg_tiles_csc_order(c, &t, &o));
    const size_t nt = (size_t)t.numtile;
    matrix->tilem = t.tilem;
    matrix->tilen = t.tilen;
    matrix->numtile = t.numtile;
    matrix->tile_ptr = (int *)fetch(c, o.tile_ptr, ((size_t)t.tilem + 1) * 4);
    matrix->tile_columnidx = (int *)fetch(c, o.tile_colidx, nt * 4);
    matrix->tile_rowidx = (int *)fetch(c, o.tile_rowidx, nt * 4);
    matrix->csc_tile_ptr = (int *)fetch(c, o.csc_ptr, ((size_t)t.tilen + 1) * 4);
    matrix->csc_tile_rowidx = (int *)fetch(c, o.csc_rowidx, nt * 4);
    matrix->tile_nnz = (int *)fetch(c, o.tile_nnz, (nt + 1) * 4);
    matrix->tile_csr_Ptr = (unsigned char *)fetch(c, o.csr_ptr, nt * 16);
    matrix->tile_csr_Col = (unsigned char *)fetch(c, o.csr_col, (size_t)t.nnz);
    matrix->tile_csr_Value = (double *)fetch(c, o.val, (size_t)t.nnz * 8);
    matrix->mask = (unsigned short *)fetch(c, o.mask, nt * 32);
    tsg_tiles_free(c, &o);
    tsg_tiles_free(c, &t);
    tsg_dcsr_free(c, &d);
}

void matrix_destroy(SMatrix *matrix) {
    free(matrix->tile_ptr);
    free(matrix->tile_columnidx);
    free(matrix->tile_rowidx);
    free(matrix->tile_nnz);
    free(matrix->tile_csr_Value);
    free(matrix->tile_csr_Col);
    free(matrix->tile_csr_Ptr);
    free(matrix->mask);
    free(matrix->csc_tile_ptr);
    free(matrix->csc_tile_rowidx);
    matrix->tile_ptr = matrix->tile_columnidx = matrix->tile_rowidx = matrix->tile_nnz = NULL;
    matrix->tile_csr_Value = NULL;
    matrix->tile_csr_Col = matrix->tile_csr_Ptr = NULL;
    matrix->mask = NULL;
    matrix->csc_tile_ptr = matrix->csc_tile_rowidx = NULL;
}

void matrix_transposition(const int m, const int n, const MAT_PTR_TYPE nnz, const MAT_PTR_TYPE *csrRowPtr,
                          const int *csrColIdx, const MAT_VAL_TYPE *csrVal, int *cscRowIdx, MAT_PTR_TYPE *cscColPtr,
                          MAT_VAL_TYPE *cscVal) {
    tsg_ctx *c = tsg_compat_ctx();
    tsg_hcsr h = {m, n, nnz, (int *)csrRowPtr, (int *)csrColIdx, (double *)csrVal};
    tsg_dcsr d, dt;
    CK("upload CSR", tsg_dcsr_upload(c, &h, &d));
    CK("transpose", tsg_dcsr_transpose(c, &d, &dt));
    CK("download", tsg_d2h(c, cscColPtr, dt.rowptr, ((size_t)n + 1) * 4));
    CK("download", tsg_d2h(c, cscRowIdx, dt.colidx, (size_t)nnz * 4));
    CK("download", tsg_d2h(c, cscVal, dt.val, (size_t)nnz * 8));
    tsg_dcsr_free(c, &dt);
    tsg_dcsr_free(c, &d);
}

/* Device tiles whose arrays are uploaded copies of an SMatrix's tile fields.
 * col_major: data arrays in CSC tile order (csr2tile_col_major output), so
 * the CSC index maps position p to data slot p. */
static void upload_tiles(tsg_ctx *c, const SMatrix *s, int col_major, tsg_dtiles *t) {
    memset(t, 0, sizeof(*t));
    const size_t nt = (size_t)s->numtile;
    t->m = s->m;
    t->n = s->n;
    t->tilem = s->tilem;
    t->tilen = s->tilen;
    t->numtile = s->numtile;
    t->nnz = s->nnz;
#define UP(field, src, bytes)                                          \
    do {                                                               \
        CK("alloc", tsg_dmalloc(c, (void **)&t->field, (bytes) + 16)); \
        CK("upload", tsg_h2d(c, t->field, (src), (bytes)));           \
    } while (0)
    UP(tile_ptr, s->tile_ptr, ((size_t)s->tilem + 1) * 4);
    UP(tile_colidx, s->tile_columnidx, nt * 4);
    UP(tile_nnz, s->tile_nnz, (nt + 1) * 4);
    UP(csr_ptr, s->tile_csr_Ptr, nt * 16);
    UP(val, s->tile_csr_Value, (size_t)s->nnz * 8);
    UP(mask, s->mask, nt * 32);
    if (col_major) {
        UP(csc_ptr, s->csc_tile_ptr, ((size_t)s->tilen + 1) * 4);
        UP(csc_rowidx, s->csc_tile_rowidx, nt * 4);
        CK("alloc", tsg_dmalloc(c, (void **)&t->csc_tileid, nt * 4 + 16));
        CK("iota", tsg_d_iota(c, t->csc_tileid, (long long)nt));
    }
    /* The matrix's own CSR (rows ascending, checked when its tiles were
     * built) is the element row index the hypersparse kernels walk. */
    if (s->rowpointer && s->columnindex && s->value) {
        UP(rowx_ptr, s->rowpointer, ((size_t)s->m + 1) * 4);
        UP(rowx_col, s->columnindex, (size_t)s->nnz * 4);
        UP(rowx_val, s->value, (size_t)s->nnz * 8);
    }
#undef UP
}

void tilespgemm(SMatrix *matrixA, SMatrix *matrixB, SMatrix *matrixC, unsigned int *blk_intersec_bitmask_A,
                unsigned int *blk_intersec_bitmask_B, int blk_intersec_bitmask_len, double densityA,
                double densityB, unsigned long long int nnzCub, unsigned long long int *nnzC_computed,
                double *compression_rate, double *time_tile, double *gflops_tile, char *filename,
                double *time_step1, double *time_step2, double *time_step3, double *time_malloc) {
    (void)blk_intersec_bitmask_A;
    (void)blk_intersec_bitmask_B;
    (void)blk_intersec_bitmask_len;
    (void)densityA;
    (void)densityB;
    (void)filename;
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dtiles A, Bcsc, B, C;
    upload_tiles(c, matrixA, 0, &A);
    upload_tiles(c, matrixB, 1, &Bcsc);
    CK("B to row-major tile order", tsg_tiles_from_csc_order(c, &Bcsc, &B));
    /* B's element row index does not depend on the tile order */
    B.rowx_ptr = Bcsc.rowx_ptr;
    B.rowx_col = Bcsc.rowx_col;
    B.rowx_val = Bcsc.rowx_val;
    Bcsc.rowx_ptr = Bcsc.rowx_col = NULL;
    Bcsc.rowx_val = NULL;
    tsg_tiles_free(c, &Bcsc);
    double tmin = 1e300;
    tsg_stats st, best;
    memset(&best, 0, sizeof(best));
    for (int ri = 0; ri < TSG_REPEAT_NUM; ri++) {
        CK("spgemm", tsg_spgemm(c, &A, &B, NULL, &C, &st));
        if (st.t_total_ms < tmin) {
            tmin = st.t_total_ms;
            best = st;
        }
        if (ri == 0) {
            printf("step1 ----Calculate the number and tile-column index of tiles of matrixC---\n");
            printf("step1 ---------------------- Runtime is  %.2f ms-------------------------\n", st.t_step1_ms);
            printf("\nstep2 --------Calculate the number of nonzeros of each tile of matrixC-----\n");
            printf("step2 ---------------------- Runtime is  %.2f ms-------------------------\n", st.t_step2_ms);
            printf("\nstep3 ---------Calculate the val&col of nonzeros of matrixC-------------\n");
            printf("step3 ---------------------- Runtime is  %.2f ms------------------------\n", st.t_step3_ms);
        }
        if (ri != TSG_REPEAT_NUM - 1) tsg_tiles_free(c, &C);
    }
    const size_t nt = (size_t)C.numtile;
    matrixC->m = matrixA->m;
    matrixC->n = matrixB->n;
    matrixC->tilem = matrixA->tilem;
    matrixC->tilen = matrixB->tilen;
    matrixC->numtile = C.numtile;
    matrixC->nnz = C.nnz;
    matrixC->tile_ptr = (int *)fetch(c, C.tile_ptr, ((size_t)C.tilem + 1) * 4);
    matrixC->tile_columnidx = (int *)fetch(c, C.tile_colidx, nt * 4);
    matrixC->tile_nnz = (int *)fetch(c, C.tile_nnz, (nt + 1) * 4);
    matrixC->tile_csr_Value = (double *)fetch(c, C.val, (size_t)C.nnz * 8);
    matrixC->tile_csr_Col = (unsigned char *)fetch(c, C.csr_col, (size_t)C.nnz);
    matrixC->tile_csr_Ptr = (unsigned char *)fetch(c, C.csr_ptr, nt * 16);
    tsg_tiles_free(c, &C);
    tsg_tiles_free(c, &A);
    tsg_tiles_free(c, &B);

    *nnzC_computed = (unsigned long long)best.nnzC;
    *compression_rate = best.nnzC ? (double)nnzCub / (double)best.nnzC : 0.0;
    *time_tile = tmin;
    *gflops_tile = tmin > 0 ? 2.0 * (double)nnzCub / (tmin * 1e6) : 0.0;
    *time_step1 = best.t_step1_ms;
    *time_step2 = best.t_step2_ms;
    *time_step3 = best.t_step3_ms;
    *time_malloc = best.t_alloc_ms;
    printf("Non-empty tiles of C = %i\n", matrixC->numtile);
    printf("nnzC = %i\n", matrixC->nnz);
    printf("HIP  TileSpGEMM runtime is %4.2f ms, gflops = %4.2f\n", tmin, *gflops_tile);
}

/* Device tile -> CSR for a C produced by tilespgemm(): the mask is not part
 * of the returned fields (src/tilespgemm-cuda.h:3141-3152), so row counts
 * come from tile_csr_Ptr as in src/tile2csr.h:8-32. */
void tile2csr(SMatrix *matrix) {
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dtiles t;
    memset(&t, 0, sizeof(t));
    const size_t nt = (size_t)matrix->numtile;
    t.m = matrix->m;
    t.n = matrix->n;
    t.tilem = matrix->tilem;
    t.tilen = matrix->tilen;
    t.numtile = matrix->numtile;
    t.nnz = matrix->nnz;
#define UP(field, src, bytes)                                         \
    do {                                                              \
        CK("alloc", tsg_dmalloc(c, (void **)&t.field, (bytes) + 16)); \
        CK("upload", tsg_h2d(c, t.field, (src), (bytes)));           \
    } while (0)
    UP(tile_ptr, matrix->tile_ptr, ((size_t)t.tilem + 1) * 4);
    UP(tile_colidx, matrix->tile_columnidx, nt * 4);
    UP(tile_nnz, matrix->tile_nnz, (nt + 1) * 4);
    UP(csr_ptr, matrix->tile_csr_Ptr, nt * 16);
    UP(csr_col, matrix->tile_csr_Col, (size_t)t.nnz);
    UP(val, matrix->tile_csr_Value, (size_t)t.nnz * 8);
#undef UP
    tsg_dcsr d;
    CK("tile2csr", tsg_tile2csr(c, &t, &d));
    matrix->rowpointer = (int *)fetch(c, d.rowptr, ((size_t)d.m + 1) * 4);
    matrix->columnindex = (int *)fetch(c, d.colidx, (size_t)d.nnz * 4);
    matrix->value = (double *)fetch(c, d.val, (size_t)d.nnz * 8);
    matrix->nnz = d.nnz;
    tsg_dcsr_free(c, &d);
    tsg_tiles_free(c, &t);
}

unsigned long long tsg_nnzcub(const SMatrix *A, const SMatrix *B) {
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dcsr da, db;
    upload_csr(c, A, &da);
    upload_csr(c, B, &db);
    unsigned long long v = 0;
    CK("nnzcub", tsg_dcsr_nnzcub(c, &da, &db, &v));
    tsg_dcsr_free(c, &da);
    tsg_dcsr_free(c, &db);
    return v;
}
