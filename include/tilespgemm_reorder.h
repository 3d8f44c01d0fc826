/*
 * tilespgemm_reorder.h -- the reorder path of the reference driver
 * (src/reorder.h, src/external_reorder.h) on the MI355X library.
 *
 * Types keep the reference layout field for field.  Orderings are computed
 * on the device (include/tsg.h: tsg_order_rcm / tsg_order_degree /
 * tsg_order_random) instead of forking reorder_algorithm/<algo>/order_gen
 * through system() and reading an order file back
 * (src/external_reorder.h:100-141).  Permutations are applied and restored
 * on the device.  Status codes follow the reference: 0 on success,
 * negative on error (nothing aborts).
 */
#ifndef TILESPGEMM_REORDER_H
#define TILESPGEMM_REORDER_H

#include "tilespgemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/reorder.h:29-34 */
typedef struct {
    int n;
    int *old_id; /* new id -> original id */
    int *new_id; /* original id -> new id */
} AxisPermutation;

/* src/reorder.h:36-63 */
typedef struct {
    int enabled;
    int n;
    AxisPermutation row_perm;
    AxisPermutation inner_perm;
    AxisPermutation col_perm;
    long long input_tiles_A_before;
    long long input_tiles_A_after;
    long long input_tiles_B_before;
    long long input_tiles_B_after;
    long long estimated_c_tiles_before;
    long long estimated_c_tiles_after;
    double reorder_time_ms;
    double signature_time_ms;
    double tile_profile_time_ms;
    double affinity_time_ms;
    double sort_time_ms;
    double local_refine_time_ms;
    double proxy_time_ms;
    double permute_time_ms;
    double risk_input_tile_ratio;
    double risk_c_tile_ratio;
    int local_refine_window;
    int proxy_sample_rows;
    int guard_degraded_to_identity;
    int structural_conservative;
    int shared_symmetric_csr;
    char mode[64];
    char risk_reason[128];
} ReorderInfo;

/* src/reorder.h:65-70 */
typedef enum {
    REORDER_PROBLEM_AA = 0,
    REORDER_PROBLEM_AAT = 1,
    REORDER_PROBLEM_AB = 2
} ReorderProblemKind;

/* src/reorder.h:72-77 */
typedef struct {
    ReorderProblemKind kind;
    const SMatrix *matrixA;
    const SMatrix *matrixB;
} ReorderProblem;

/* src/reorder.h:129-190 */
void axis_permutation_init(AxisPermutation *perm);
void axis_permutation_destroy(AxisPermutation *perm);
int axis_permutation_alloc(AxisPermutation *perm, int n);
int axis_permutation_identity(AxisPermutation *perm, int n);
int axis_permutation_copy(AxisPermutation *dst, const AxisPermutation *src);
int axis_permutation_is_identity(const AxisPermutation *perm);
/* src/reorder.h:201-238 */
void reorder_info_init(ReorderInfo *info);
void reorder_info_destroy(ReorderInfo *info);

/* Reads a new->old order file (one zero-based id per line, whitespace
 * separated) of n ids into perm (both maps).  Returns 0, -10 cannot open,
 * -11 fewer than n ids, -12 id outside [0, n), -13 duplicate id, -14 more
 * than n ids (messages as the reference prints them).
 * Replaces: read_external_order_file  src/external_reorder.h:53-98 */
int read_external_order_file(const char *path, int n, AxisPermutation *perm);

/* Ordering plan for an AA / AAT / AB problem with the axis sharing of
 * build_external_reorder_plan (src/external_reorder.h:204-319): AA and AAT
 * use A's order on all three axes, AB uses A's order for rows and the
 * inner axis and B's order for columns.  `algorithm` "rcm", "degree"
 * (ids by descending row length, ties by id), "random" (seed from
 * TSG_RANDOM_SEED, default 1) and "original" are computed on the device
 * from the matrices themselves (the paths are then not read and may be
 * NULL); any other name runs the reference's adapter
 * $TASA_REORDER_ALGORITHM_ROOT/<algorithm>/order_gen <path> <order file>
 * (root default "reorder_algorithm", order files under
 * $TASA_REORDER_ORDER_DIR, default /tmp/tasa_reorder_orders) on
 * matrix_path_A, and for AB on matrix_path_B (or A's path), and reads its
 * order file back.  Status codes are the reference's: -1 bad arguments,
 * -2 missing B, -3/-8 inner dimensions differ, -4/-5/-6/-7 non-square
 * operand, -9 unknown problem kind, -10..-14 bad order file (see
 * read_external_order_file), -20 adapter missing or failed (also: a
 * device order with no device), -30 device order failed.  Every argument
 * check happens before any device work.  The tile estimates are exact
 * device counts, or -1 when no device is present.
 * Replaces: build_external_reorder_plan  src/external_reorder.h:204 */
int build_external_reorder_plan(ReorderProblemKind kind, const char *algorithm, const char *matrix_path_A,
                                const char *matrix_path_B, const SMatrix *matrixA, const SMatrix *matrixB,
                                ReorderInfo *info, double *ordergen_ms);

/* Identity plan flagged as degraded to identity.
 * Replaces: build_original_reorder_plan  src/external_reorder.h:173 */
int build_original_reorder_plan(ReorderProblemKind kind, const SMatrix *matrixA, const SMatrix *matrixB,
                                ReorderInfo *info);

/* The reference's TASA/GTSP heuristic planner is not implemented here;
 * this entry point builds the RCM plan of build_external_reorder_plan for
 * the problem instead (mode "rcm").  Codes as the reference: -1 NULL
 * argument, -3 missing B for AAT/AB, -2 unknown kind.
 * Replaces: int build_reorder_plan(...)  src/reorder.h:2407 */
int build_reorder_plan(const ReorderProblem *problem, ReorderInfo *info);

/* output = P*input*Q^T (row r <- input row row_perm->old_id[r], column c ->
 * col_perm->new_id[c], rows sorted), computed on the device; output arrays
 * malloc'd, output->isSymmetric = 0.  -1 on shape mismatch, -2 on a device
 * failure.
 * Replaces: apply_bipermutation_csr  src/reorder.h:2429 */
int apply_bipermutation_csr(const SMatrix *input, const AxisPermutation *row_perm,
                            const AxisPermutation *col_perm, SMatrix *output);

/* In-place inverse of apply for a result C: row r -> row_perm->old_id[r],
 * column c -> col_perm->old_id[c], rows sorted; NULL perms are a no-op.
 * The matrix's CSR arrays are freed and replaced by new malloc'd ones.
 * -1 on shape mismatch or device failure.
 * Replaces: restore_bipermutation_csr  src/reorder.h:2506 */
int restore_bipermutation_csr(SMatrix *matrix, const AxisPermutation *row_perm,
                              const AxisPermutation *col_perm);

#ifdef __cplusplus
}
#endif
#endif
