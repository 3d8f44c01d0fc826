/*
 * tsg.h -- device-resident C ABI of the MI355X TileSpGEMM library.
 *
 * The reference keeps every operand on the host and copies it to the GPU
 * inside tilespgemm() (src/tilespgemm-cuda.h:2646-2712).  This interface is
 * the device-resident layer underneath the drop-in entry points of
 * tilespgemm.h: matrices live in HBM between calls, conversions run on the
 * device, and the SpGEMM pipeline is steps 1-3 of src/tilespgemm-cuda.h:
 * 2740-3124 re-designed for wave64.  All functions return an int status
 * (TSG_OK = 0, negative on error, message in tsg_last_error()) and never
 * abort; all device memory is owned by the context's caching allocator.
 *
 * Plain C types only (ints, doubles, pointers), so ctypes / cgo / JNI can
 * bind it directly.
 */
#ifndef TSG_H
#define TSG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TSG_OK = 0,
    TSG_ERR_HIP = -1,      /* a HIP runtime call failed */
    TSG_ERR_ARG = -2,      /* invalid argument / shape mismatch */
    TSG_ERR_OVERFLOW = -3, /* a count does not fit the int32 fields */
    TSG_ERR_NOMEM = -4,    /* device allocation failed */
    TSG_ERR_BUDGET = -5,   /* request exceeds the admission budget */
    TSG_ERR_STATE = -6     /* object used in the wrong state */
};

typedef struct tsg_ctx tsg_ctx;

/* Context: one device, one HIP stream, a caching device allocator.
 * Not thread-safe: one host thread per context. */
int tsg_ctx_create(int device, tsg_ctx **out);
void tsg_ctx_destroy(tsg_ctx *ctx);
int tsg_ctx_sync(tsg_ctx *ctx);
void *tsg_ctx_stream(tsg_ctx *ctx);           /* the hipStream_t */
long long tsg_ctx_bytes_in_use(tsg_ctx *ctx); /* live allocations */
long long tsg_ctx_bytes_cached(tsg_ctx *ctx); /* cached free blocks */
void tsg_ctx_trim(tsg_ctx *ctx);              /* return cached blocks */
const char *tsg_last_error(void);             /* per host thread */

/* ------------------------------------------------------------------ */
/* Host CSR (malloc'd arrays, int32 row pointers as MAT_PTR_TYPE).    */
typedef struct {
    int m, n, nnz;
    int *rowptr;    /* m+1 */
    int *colidx;    /* nnz, ascending inside each row */
    double *val;    /* nnz */
} tsg_hcsr;
void tsg_hcsr_free(tsg_hcsr *h);

/* Value modes for generators and tsg_hcsr_set_values. */
enum {
    TSG_VAL_KEEP = 0,     /* generator default (stencil constants) */
    TSG_VAL_UNIFORM_PM1 = 1, /* U[-1,1) from (seed, entry index) */
    TSG_VAL_UNIFORM_01 = 2,  /* U(0,1] from (seed, entry index) */
    TSG_VAL_MOD10 = 3        /* value[i] = i % 10, src/main.cu:191-192 */
};
int tsg_hcsr_set_values(tsg_hcsr *h, int mode, unsigned long long seed);
/* In place: each row stably sorted by column, duplicate entries summed in
 * stored order (nnz shrinks; arrays are not reallocated). */
int tsg_hcsr_canonicalize(tsg_hcsr *h);
/* Matrix Market file -> canonical CSR: mmio_allinone (tilespgemm.h), then
 * tsg_hcsr_canonicalize.  *is_symmetric = the header's symmetric flag. */
int tsg_read_mtx(const char *path, tsg_hcsr *out, int *is_symmetric);

/* 2D 5-point Laplacian on nx*ny (row = x + nx*y), diag 4, off-diag -1. */
int tsg_gen_stencil5(int nx, int ny, tsg_hcsr *out);
/* 3D 27-point stencil on nx*ny*nz (row = x + nx*y + nx*ny*z), periodic or
 * Dirichlet; values by `valmode` (TSG_VAL_KEEP gives 26 / -1). */
int tsg_gen_stencil27(int nx, int ny, int nz, int periodic, int valmode,
                      unsigned long long seed, tsg_hcsr *out);
/* Erdos-Renyi: each row draws Poisson(mean_row) distinct uniform columns. */
int tsg_gen_er(int n, double mean_row, int valmode, unsigned long long seed,
               tsg_hcsr *out);
/* Banded: row i holds every column in [i-halfbw, i+halfbw] inside [0,n). */
int tsg_gen_banded(int n, int halfbw, int valmode, unsigned long long seed,
                   tsg_hcsr *out);
/* R-MAT: 2^scale vertices, edge_factor*2^scale edges drawn with quadrant
 * probabilities (a,b,c,1-a-b-c); optional symmetrization; duplicates
 * merged; square CSR with sorted rows. */
int tsg_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 int symmetrize, int valmode, unsigned long long seed,
                 tsg_hcsr *out);

/* ------------------------------------------------------------------ */
/* Device CSR.                                                        */
typedef struct {
    int m, n, nnz;
    int *rowptr;    /* device, m+1 */
    int *colidx;    /* device, nnz */
    double *val;    /* device, nnz */
} tsg_dcsr;

int tsg_dcsr_upload(tsg_ctx *ctx, const tsg_hcsr *h, tsg_dcsr *out);
int tsg_dcsr_download(tsg_ctx *ctx, const tsg_dcsr *d, tsg_hcsr *out);
void tsg_dcsr_free(tsg_ctx *ctx, tsg_dcsr *d);
/* Stable transpose on the device (rows ascending inside each column). */
int tsg_dcsr_transpose(tsg_ctx *ctx, const tsg_dcsr *a, tsg_dcsr *at);
/* sum over nonzeros A(i,k) of nnz(B row k) (src/main.cu:391-397). */
int tsg_dcsr_nnzcub(tsg_ctx *ctx, const tsg_dcsr *a, const tsg_dcsr *b,
                    unsigned long long *out);

/* ------------------------------------------------------------------ */
/* Device tile matrix.  Data arrays are in row-major tile order; the
 * optional CSC tile index lists, for each tile column, its tiles ordered
 * by tile row (csc_tileid = row-major tile id).  csr_col holds plain
 * in-tile columns (0..15).                                             */
typedef struct {
    int m, n, tilem, tilen;
    int numtile, nnz;
    int *tile_ptr;          /* tilem+1 */
    int *tile_colidx;       /* numtile */
    int *tile_rowidx;       /* numtile */
    int *tile_nnz;          /* numtile+1, exclusive offsets */
    unsigned char *csr_ptr; /* 16*numtile */
    unsigned char *csr_col; /* nnz */
    double *val;            /* nnz */
    unsigned short *mask;   /* 16*numtile, bit (15-col) */
    int *csc_ptr;           /* tilen+1   (NULL until built) */
    int *csc_rowidx;        /* numtile */
    int *csc_tileid;        /* numtile */
    int *rowx_ptr;          /* m+1   row index (NULL until built) */
    int *rowx_col;          /* nnz: global columns, row by row ascending */
    double *rowx_val;       /* nnz */
} tsg_dtiles;

int tsg_csr2tile(tsg_ctx *ctx, const tsg_dcsr *a, tsg_dtiles *out);
int tsg_tiles_build_csc(tsg_ctx *ctx, tsg_dtiles *t);
/* Element row index of a B operand (its rows as CSR).  With it, the
 * numeric step handles tile-rows too large or too sparse for the
 * row-wise tile kernel element by element instead of by tile
 * intersection.  Optional; 12 bytes per nonzero. */
int tsg_tiles_build_rowindex(tsg_ctx *ctx, tsg_dtiles *t);
int tsg_tile2csr(tsg_ctx *ctx, const tsg_dtiles *t, tsg_dcsr *out);
void tsg_tiles_free(tsg_ctx *ctx, tsg_dtiles *t);

/* Host copy of a device tile matrix (malloc'd; NULL pointers skipped). */
typedef struct {
    int m, n, tilem, tilen, numtile, nnz;
    int *tile_ptr, *tile_colidx, *tile_rowidx, *tile_nnz;
    unsigned char *csr_ptr, *csr_col;
    double *val;
    unsigned short *mask;
    int *csc_ptr, *csc_rowidx, *csc_tileid;
} tsg_htiles;
int tsg_tiles_download(tsg_ctx *ctx, const tsg_dtiles *d, tsg_htiles *out);
void tsg_htiles_free(tsg_htiles *h);

/* Checksums of a C tile matrix, or of a tile-row view of one, so products
 * too large to download can be compared: `entries` sums a 64-bit mix of
 * (row, column, value bits) over C's nonzeros as tile2csr emits them,
 * `tiles` a mix of (tile-row, tile-column, tile nonzeros, row masks, row
 * offsets) over its tiles.  The sums are order-free, so the checksums of
 * the tile-row chunks of a product add up (mod 2^64) to those of the
 * whole.  Rows are numbered from row_base (a multiple of 16: the view's
 * first row in the whole product), tile columns relative to jbase.
 * `entries_t` is `entries` with each entry's row and column swapped: the
 * two are equal when C is bit-symmetric (A*A^T).                       */
typedef struct {
    long long numtile, nnz;
    unsigned long long tiles, entries, entries_t;
} tsg_csum;
int tsg_tiles_checksum(tsg_ctx *ctx, const tsg_dtiles *c, long long row_base, int jbase, tsg_csum *out);
/* Rows rows[0..nrows) of a C tile matrix or view (ids local to it, any
 * order, repeats allowed) as a host CSR of nrows rows (malloc'd; free with
 * tsg_hcsr_free); columns relative to tile column jbase. */
int tsg_tiles_rows(tsg_ctx *ctx, const tsg_dtiles *c, int jbase, const int *rows, int nrows, tsg_hcsr *out);

/* ------------------------------------------------------------------ */
/* SpGEMM C = A*B on device tiles.  B must carry its CSC tile index
 * (tsg_tiles_build_csc).  Passing the same object as A and B computes A^2
 * without a second copy of the operand.                               */
typedef struct {
    int dense_acc_min_nnz; /* C tiles with nnz >= this use the dense 16x16
                              LDS accumulator, smaller ones the compact
                              sparse one (default 64) */
    int s1_window_words;   /* step-1 LDS bitmap window, 32-bit words per
                              wave: 256, 1024, 2048, or 0 = by B's width
                              (default 0) */
    int force_tile_parallel; /* 1: steps 2 and 3 use the tile-parallel
                              kernels for every tile-row instead of the
                              row-wise LDS kernels (default 0 = by size) */
    int batch_group_nnz_cap; /* batches: most C nonzeros per fused group
                              (0 = the int32 limit of the tile offsets) */
    int element_mode;      /* numeric step for tile-rows the row-wise tile
                              kernel cannot hold, when B has a row index:
                              0 = element-wise (default), 1 = tile-parallel,
                              2 = element-wise for every tile-row */
    int no_heavy_rows;     /* 1: rows with very many products stay in the
                              element-wise segments instead of one wave each
                              (default 0) */
    int heavy_min_products; /* a row with more intermediate products (or more
                              than 2048 A entries) is heavy (0 = 65536) */
    int heavy_item_products; /* a heavy row is split into column parts of
                              4096 columns, handed out in runs of about this
                              many products (0 = 32768) */
} tsg_tune;
void tsg_tune_default(tsg_tune *t);

typedef struct {
    double t_step1_ms, t_step2_ms, t_step3_ms, t_alloc_ms, t_total_ms;
    long long numtileC, nnzC;
    long long tiles_sparse_acc, tiles_dense_acc, tiles_empty;
    double t_step3_rows_ms; /* the row-wise numeric kernel's launch alone */
    long long heavy_rows;   /* rows computed one wave each (k_step3_heavy) */
    long long heavy_bytes;  /* transient device bytes of that path */
} tsg_stats;

/* tune and stats may be NULL.  When stats is given, each step is bracketed
 * by HIP events on the context stream. */
int tsg_spgemm(tsg_ctx *ctx, const tsg_dtiles *a, const tsg_dtiles *b,
               const tsg_tune *tune, tsg_dtiles *c, tsg_stats *stats);
/* Step 1 alone: number of C tiles (structurally empty ones included). */
int tsg_spgemm_count_tiles(tsg_ctx *ctx, const tsg_dtiles *a, const tsg_dtiles *b, long long *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the C tile count of A*B exactly as
 * tsg_spgemm_count_tiles computes it, enqueued on ctx's stream with no
 * allocation, no memset and no host synchronization; the int64 count is
 * written to d_out, device memory the caller owns.  Its scratch is a fixed
 * TSG_NOALLOC_SCRATCH_BYTES device array per GPU, held until the process
 * ends, so A may have at most TSG_NOALLOC_MAX_TILE_ROWS tile-rows (m up to
 * about 67 million); larger inputs return TSG_ERR_ARG before any device
 * work.  Every call on one GPU shares that array: calls that overlap on a
 * GPU (several threads, or several streams) race. */
#define TSG_NOALLOC_SCRATCH_BYTES (16u << 20)
#define TSG_NOALLOC_MAX_TILE_ROWS 4192256
int tsg_spgemm_count_tiles_noalloc(tsg_ctx *ctx, const tsg_dtiles *a, const tsg_dtiles *b, long long *d_out);

/* Products whose C exceeds the int32 offsets of the tile format (or the
 * memory at hand) run in chunks of A tile-rows.  The reference has no
 * chunking: tilespgemm() computes the whole C at once
 * (src/tilespgemm-cuda.h:2740-3124) with int offsets (src/common.h:23-25).
 * tsg_spgemm_chunk_plan: tile-row bounds (bounds[0] = 0 < ... <
 * bounds[*nchunks] = tilem, at most max_chunks chunks) so that each chunk
 * has at most max_products intermediate products (an upper bound on its C
 * nonzeros); products_total (may be NULL) = nnzCub.
 * tsg_spgemm_tile_rows: C rows [16*r0, min(16*r1, m)) of A*B as a tile
 * matrix of r1-r0 tile-rows, bit-identical to those rows of the full C. */
int tsg_spgemm_chunk_plan(tsg_ctx *ctx, const tsg_dtiles *a, const tsg_dtiles *b, long long max_products,
                          int max_chunks, int *bounds, int *nchunks, unsigned long long *products_total);
int tsg_spgemm_tile_rows(tsg_ctx *ctx, const tsg_dtiles *a, const tsg_dtiles *b, int r0, int r1,
                         const tsg_tune *tune, tsg_dtiles *c, tsg_stats *stats);

/* ------------------------------------------------------------------ */
/* Raw device memory helpers for host layers (all on the context stream;
 * tsg_h2d / tsg_d2h return after the copy completed).                  */
int tsg_dmalloc(tsg_ctx *ctx, void **p, size_t bytes);
void tsg_dfree(tsg_ctx *ctx, void *p);
int tsg_h2d(tsg_ctx *ctx, void *d, const void *h, size_t bytes);
int tsg_d2h(tsg_ctx *ctx, void *h, const void *d, size_t bytes);
int tsg_d_iota(tsg_ctx *ctx, int *d, long long n);
/* Device-to-device copy by a plain streaming kernel (k_copy16); 16-byte
 * aligned pointers and size.  Used to calibrate byte counters. */
int tsg_copy_kernel(tsg_ctx *ctx, void *dst, const void *src, size_t bytes);
/* In-tile column bytes packed as (row<<4)|col (A layout of
 * src/csr2tile.h:185) into d_out[nnz]. */
int tsg_tiles_packed_cols(tsg_ctx *ctx, const tsg_dtiles *t, unsigned char *d_out);
/* Copy of t with tile_nnz/csr_ptr/csr_col/val/mask permuted into CSC tile
 * order (B layout of src/csr2tile.h:327-477); csc_tileid = identity and the
 * row-major tile index kept for step 1.  Builds t's CSC index if needed. */
int tsg_tiles_csc_order(tsg_ctx *ctx, tsg_dtiles *t, tsg_dtiles *out);
/* Inverse: `in` holds tile data in CSC tile order plus both indices (the
 * csr2tile_col_major layout); out gets the data in row-major tile order
 * with csc_tileid mapping each CSC position to its row-major tile. */
int tsg_tiles_from_csc_order(tsg_ctx *ctx, const tsg_dtiles *in, tsg_dtiles *out);

/* ------------------------------------------------------------------ */
/* Orderings and permutations on the device (the reorder path of
 * src/reorder.h / src/external_reorder.h / reorder_algorithm/).  A
 * permutation is given as old_id[new] (device int array of n entries);
 * tsg_perm_invert gives new_id[old].                                  */
int tsg_perm_validate(tsg_ctx *ctx, int n, const int *d_perm);
int tsg_perm_invert(tsg_ctx *ctx, int n, const int *d_old_id, int *d_new_id);
/* ids sorted by (-degree, id); degree = row length, plus the mirrored
 * count when the matrix stands for a symmetric-header file
 * (reorder_algorithm/common/order_utils.py:41-45,78-81). */
int tsg_order_degree(tsg_ctx *ctx, const tsg_dcsr *a, int symmetric_header, int *d_old_id);
/* Reverse Cuthill-McKee on pattern(A + A^T) (self loops kept; degree =
 * row length + 1 if the diagonal is stored): seeds and children in
 * ascending (degree, id).  levels_out (may be NULL) = BFS depth + 1. */
int tsg_order_rcm(tsg_ctx *ctx, const tsg_dcsr *a, int *d_old_id, int *levels_out);
/* Seeded uniform random order (keys from a 64-bit mix of (seed, id)). */
int tsg_order_random(tsg_ctx *ctx, int n, unsigned long long seed, int *d_old_id);
/* pattern(A + A^T) with sorted rows (values not set). */
int tsg_dcsr_symmetrize(tsg_ctx *ctx, const tsg_dcsr *a, tsg_dcsr *g);
/* out = P*A*Q^T: out row r is row d_row_old[r] of A, column c becomes
 * d_col_new[c], rows sorted by column (src/reorder.h:2429-2504).  With
 * row map = new_id and column map = old_id it is the restore of
 * src/reorder.h:2506-2551. */
int tsg_dcsr_permute(tsg_ctx *ctx, const tsg_dcsr *a, const int *d_row_old, const int *d_col_new,
                     tsg_dcsr *out);

/* ------------------------------------------------------------------ */
/* Batched SpGEMM: many independent products per step.  There is no
 * counterpart in the reference (one product per tilespgemm() call,
 * src/tilespgemm-cuda.h:2610); each product is the one tilespgemm()
 * computes for the driver's A^2 or AA^T mode (src/main.cu:199-241), and
 * each result is bit-identical to tsg_spgemm on the same operands.
 * The problems are packed block-diagonally on 16-row/column boundaries so
 * every pipeline kernel runs once for the whole batch; the pipeline is
 * planned once and captured in a HIP graph, and tsg_batch_run is one
 * graph launch recomputing every product into resident C buffers.      */
enum { TSG_OP_A2 = 0, TSG_OP_AAT = 1 };
typedef struct tsg_batch tsg_batch;
typedef struct {
    int count;            /* problems */
    int groups;           /* fused pipelines (split where C overflows int32) */
    long long numtileC;   /* C tiles over all problems */
    long long nnzC;       /* C nonzeros over all problems */
    int graph_nodes;
    double build_ms;      /* packing, tile conversion and planning */
    double capture_ms;    /* graph capture + instantiation */
    tsg_tune tune;        /* configuration in use */
} tsg_batch_info;
/* problems[i] are device CSRs (read during create only; the batch keeps
 * its own packed copy); ops[i] is TSG_OP_A2 (square) or TSG_OP_AAT. */
int tsg_batch_create(tsg_ctx *ctx, const tsg_dcsr *problems, const int *ops, int count,
                     const tsg_tune *tune, tsg_batch **out);
/* One step: enqueue the captured graph on the context stream (async). */
int tsg_batch_run(tsg_batch *b);
/* Times every candidate configuration over `reps` graph launches and
 * keeps the fastest (chosen / best_ms may be NULL). */
int tsg_batch_autotune(tsg_batch *b, int reps, tsg_tune *chosen, double *best_ms);
int tsg_batch_get_info(const tsg_batch *b, tsg_batch_info *out);
/* Problem i's C (after a run): tile arrays in the problem's own indices
 * and/or CSR; either pointer may be NULL.  Synchronizes the stream. */
int tsg_batch_result(tsg_batch *b, int i, tsg_htiles *tiles, tsg_hcsr *csr);
/* Checksums (tsg_tiles_checksum) of problem i's C in its own row and
 * column numbering, computed on the device.  Synchronizes the stream. */
int tsg_batch_result_checksum(tsg_batch *b, int i, tsg_csum *out);
void tsg_batch_destroy(tsg_batch *b);

/* The batched workload: `count` problems from `seed`; n log-uniform in
 * [nmin, nmax]; family 0 = Erdos-Renyi with mean nnz/row U[4,32], 1 =
 * banded with half-bandwidth uniform in {2..40}; op A^2 or AA^T with
 * probability 1/2 each; seed_i for the generator of problem i. */
typedef struct {
    int n, family, op, halfbw;
    double mean_row;
    unsigned long long seed;
} tsg_batch_spec;
int tsg_gen_batch_spec(int count, unsigned long long seed, int nmin, int nmax, tsg_batch_spec *out);
/* Generates problem `s` (values U[-1,1)). */
int tsg_gen_batch_problem(const tsg_batch_spec *s, tsg_hcsr *out);

/* ------------------------------------------------------------------ */
/* Serving: continuous batching of A^2 / AA^T requests on resident
 * matrices.  There is no counterpart in the reference (its driver runs one
 * product per process, src/main.cu:15-638).  Registration uploads a matrix
 * once and keeps its CSR, tiles (with CSC and row indices) and those of
 * A^T in HBM; it also runs each product once to size it exactly, and that
 * symbolic size (C's tiles and nonzeros, plus the pipeline's work) is the
 * request's admission estimate.  Worker threads (slots, one HIP stream
 * each) take the first queued request that fits the budget left, fused
 * with later small ones into one block-diagonal product; every result is
 * bit-identical to tsg_spgemm on the same operands and stays in HBM until
 * released.  All calls are thread-safe.                               */
typedef struct tsg_server tsg_server;
typedef struct {
    long long budget_bytes;   /* HBM for resident operands + admitted requests (64 GiB) */
    int slots;                /* worker streams (4) */
    int max_batch;            /* requests fused into one product (64) */
    long long small_bytes;    /* requests at most this large are fused (256 MiB) */
    long long chunk_products; /* tile-row chunk cap on intermediate products (1.5e9) */
    double starve_ms;         /* a head request waiting this long stops smaller
                                 ones from passing it (200) */
    tsg_tune tune;
} tsg_server_config;
void tsg_server_config_default(tsg_server_config *c);

typedef struct {
    int status, handle, op;
    int chunks;               /* tile-row pieces of C */
    int fused_with;           /* requests in the same fused product (1 = alone) */
    long long numtileC, nnzC, bytes;
    double t_arrival_ms, t_admit_ms, t_done_ms; /* server clock */
    double t_gpu_ms;          /* device time of its pipeline (of the fused
                                 product when fused) */
} tsg_request_info;

typedef struct {
    long long resident_bytes, reserved_bytes, peak_reserved_bytes;
    int queued;
    long long fused_launches, single_launches;
    long long cached_bytes;         /* free blocks held by the workers' allocators */
    long long peak_footprint_bytes; /* most of resident + reserved + cached, which
                                       bounds the server's device memory and stays
                                       within the budget */
} tsg_server_stats_t;

int tsg_server_create(int device, const tsg_server_config *cfg, tsg_server **out);
void tsg_server_destroy(tsg_server *s);
int tsg_server_register(tsg_server *s, const tsg_hcsr *a, int *handle);
/* Exact size of a registered product (any pointer may be NULL). */
int tsg_server_size(tsg_server *s, int handle, int op, long long *numtileC, long long *nnzC,
                    long long *bytes, int *chunks);
int tsg_server_submit(tsg_server *s, int handle, int op, unsigned long long *id);
/* Blocks until request id is done (timeout_ms <= 0: no limit); returns its status. */
int tsg_server_wait(tsg_server *s, unsigned long long id, double timeout_ms, tsg_request_info *info);
int tsg_server_result_csr(tsg_server *s, unsigned long long id, tsg_hcsr *out);
/* Checksums (tsg_tiles_checksum) of a finished request's C, summed over
 * its chunks or taken on its share of a fused product. */
int tsg_server_result_checksum(tsg_server *s, unsigned long long id, tsg_csum *out);
/* Rows rows[0..nrows) of a finished request's C as a host CSR. */
int tsg_server_result_rows(tsg_server *s, unsigned long long id, const int *rows, int nrows, tsg_hcsr *out);
int tsg_server_release(tsg_server *s, unsigned long long id);
int tsg_server_stats(tsg_server *s, tsg_server_stats_t *out);
/* Open-loop load: request i submitted at arrival_ms[i] after the call;
 * returns when all are done (infos may be NULL).  keep = 0 releases each
 * result on completion. */
int tsg_server_replay(tsg_server *s, int n, const int *handles, const int *ops, const double *arrival_ms,
                      int keep, unsigned long long *ids, tsg_request_info *infos);
/* The same with progress(submitted, completed, user) called from the
 * calling thread after every submission (completed = requests already
 * done) and, once all are submitted, after each request it has waited
 * for (completed = requests waited for). */
typedef void (*tsg_replay_progress)(int submitted, int completed, void *user);
int tsg_server_replay_progress(tsg_server *s, int n, const int *handles, const int *ops, const double *arrival_ms,
                               int keep, unsigned long long *ids, tsg_request_info *infos,
                               tsg_replay_progress progress, void *user);

/* Library build information (compile flags, target). */
const char *tsg_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* TSG_H */
