/*
 * tilespgemm.h -- drop-in C ABI of the TileSpGEMM hot path (C = A*A, C = A*A^T)
 * for AMD Instinct MI355X (gfx950).
 *
 * Every declaration here replaces one entry point of the reference driver's
 * headers (cited file:line, paths relative to the reference repository root).
 * The host structures keep the reference layout field for field, so a driver
 * compiled against the reference `SMatrix` links against this library
 * unchanged.  All device work runs in hand-written HIP kernels; the host
 * functions below only move arrays between host and device.
 *
 * Tile format (16x16 tiles, BLOCK_SIZE = 16), as produced by
 * csr2tile_row_major (src/csr2tile.h:196-259):
 *   tile_ptr[tilem+1]       tile offsets per tile-row
 *   tile_columnidx[T]       tile column, ascending inside a tile-row
 *   tile_rowidx[T]          tile row
 *   tile_nnz[T+1]           exclusive nonzero offsets per tile
 *   tile_csr_Ptr[16*T]      per-row start inside the tile (u8); rows past the
 *                           matrix edge hold the tile's nnz
 *   tile_csr_Col[nnz]       A (row-major): (row<<4)|col; B (col-major): col
 *   tile_csr_Value[nnz]     FP64 values, row-major inside a tile, ascending col
 *   mask[16*T]              row bitmaps, bit (15-col) set for each nonzero
 * csr2tile_col_major (src/csr2tile.h:261-478) keeps tile_ptr/tile_columnidx in
 * row-major tile order and stores tile_nnz/tile_csr_xxx/mask in column-major
 * tile order, indexed by csc_tile_ptr[tilen+1] / csc_tile_rowidx[T].
 */
#ifndef TILESPGEMM_H
#define TILESPGEMM_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MAT_VAL_TYPE
#define MAT_VAL_TYPE double
#endif
#ifndef MAT_PTR_TYPE
#define MAT_PTR_TYPE int
#endif

#define TSG_BLOCK_SIZE 16
/* Timed repeats per tilespgemm() call; the minimum is reported
 * (src/common.h:69, src/tilespgemm-cuda.h:3177-3180). */
#define TSG_REPEAT_NUM 5

#ifndef SMATRIX
#define SMATRIX
/* Same fields, order and types as src/common.h:113-135. */
typedef struct
{
    int m;
    int n;
    int nnz;
    int isSymmetric;
    MAT_VAL_TYPE *value;
    int *columnindex;
    MAT_PTR_TYPE *rowpointer;
    int tilem;
    int tilen;
    MAT_PTR_TYPE *tile_ptr;
    int *tile_columnidx;
    int *tile_rowidx;
    int *tile_nnz;
    int numtile;
    MAT_VAL_TYPE *tile_csr_Value;
    unsigned char *tile_csr_Col;
    unsigned char *tile_csr_Ptr;
    unsigned short *mask;
    int *csc_tile_ptr;
    int *csc_tile_rowidx;
} SMatrix;
#endif

/* CSR -> row-major tiles, computed on the device; fills every tile field of
 * `matrix` with malloc'd host arrays (tile_rowidx included).
 * Replaces: void csr2tile_row_major(SMatrix *matrix)  src/csr2tile.h:196 */
void csr2tile_row_major(SMatrix *matrix);

/* CSR -> column-major tiles (B operand layout), computed on the device.
 * Replaces: void csr2tile_col_major(SMatrix *matrix)  src/csr2tile.h:261 */
void csr2tile_col_major(SMatrix *matrix);

/* Tile -> CSR (sorted columns, explicit zeros kept), computed on the device;
 * mallocs rowpointer/columnindex/value and sets nnz.
 * Replaces: void tile2csr(SMatrix *matrix)  src/tile2csr.h:71 */
void tile2csr(SMatrix *matrix);

/* Frees the tile arrays csr2tile_* allocated (tile_rowidx and the CSC index
 * included, unlike the reference which leaks them).
 * Replaces: void matrix_destroy(SMatrix *matrix)  src/csr2tile.h:481 */
void matrix_destroy(SMatrix *matrix);

/* Stable CSR -> CSC transpose computed on the device (rows ascending inside
 * each column).  cscColPtr has n+1 entries.
 * Replaces: void matrix_transposition(...)  src/utils.h:158-195 */
void matrix_transposition(const int m, const int n, const MAT_PTR_TYPE nnz,
                          const MAT_PTR_TYPE *csrRowPtr, const int *csrColIdx,
                          const MAT_VAL_TYPE *csrVal, int *cscRowIdx,
                          MAT_PTR_TYPE *cscColPtr, MAT_VAL_TYPE *cscVal);

/* C = A*B with A from csr2tile_row_major and B from csr2tile_col_major.
 * Uploads A and B, runs the device pipeline (step 1 tile symbolic, step 2
 * bitmask symbolic, step 3 numeric) TSG_REPEAT_NUM times timed with HIP
 * events, reports the minimum, and fills matrixC's tile fields (tile_ptr,
 * tile_columnidx, tile_nnz, numtile, nnz, m, n, tilem, tilen, tile_csr_Value,
 * tile_csr_Col = plain column, tile_csr_Ptr) with malloc'd host arrays, as
 * src/tilespgemm-cuda.h:3141-3152 does.  The intersection bitmasks and
 * densities are accepted for signature compatibility and not needed (NULL is
 * fine).  On a HIP error or an int32 overflow of nnz(C) it prints the reason
 * and aborts, since the reference entry point returns void.
 * Replaces: void tilespgemm(...)  src/tilespgemm-cuda.h:2610-2624 */
void tilespgemm(SMatrix *matrixA, SMatrix *matrixB, SMatrix *matrixC,
                unsigned int *blk_intersec_bitmask_A,
                unsigned int *blk_intersec_bitmask_B,
                int blk_intersec_bitmask_len,
                double densityA, double densityB,
                unsigned long long int nnzCub,
                unsigned long long int *nnzC_computed,
                double *compression_rate, double *time_tile,
                double *gflops_tile, char *filename,
                double *time_step1, double *time_step2, double *time_step3,
                double *time_malloc);

/* Number of intermediate products: sum over A's nonzeros of the length of the
 * B row they select (src/main.cu:391-397).  Computed on the device. */
unsigned long long tsg_nnzcub(const SMatrix *A, const SMatrix *B);

/* Matrix Market coordinate file -> CSR with the reference loader's
 * semantics: symmetric/hermitian headers expanded to both triangles,
 * pattern -> 1.0, duplicates kept, each row in file order (so unsorted
 * unless the file is stored column-major); arrays malloc'd.  Returns 0, or
 * -1 cannot open, -2 bad banner, -4 bad size line, -5 entry out of range
 * or file too short.  Host only (no device work).
 * Replaces: int mmio_allinone(...)  src/mmio_highlevel.h:593-759 */
int mmio_allinone(int *m, int *n, MAT_PTR_TYPE *nnz, int *isSymmetric,
                  MAT_PTR_TYPE **csrRowPtr, int **csrColIdx,
                  MAT_VAL_TYPE **csrVal, char *filename);

/* The reference driver, `tsg_driver -d <dev> -aat {0|1} <A> [-b <B>]
 * [-reorder 1] [-reorder_algo rcm|degree|random|original|tasa|<adapter>]`
 * with the flow of src/main.cu:15-638 (load, value override i % 10, B = A,
 * A^T or the -b file, optional reorder plan + apply, nnzCub, csr2tile,
 * tilespgemm, tile2csr, restore, check).  Returns the process exit code.
 * Replaces: int main(int argc, char **argv)  src/main.cu:15 */
int tsg_driver_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif
#endif /* TILESPGEMM_H */
