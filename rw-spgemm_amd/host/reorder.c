/*
 * reorder.c -- the reference's reorder API (include/tilespgemm_reorder.h)
 * in C on top of the device ordering / permutation kernels.
 *
 * Plans follow the axis-sharing rules of build_external_reorder_plan
 * (src/external_reorder.h:204-319); the order itself is computed on the
 * device rather than by an external order_gen process.  The before/after
 * tile statistics the reference estimates from sampled rows
 * (src/external_reorder.h:143-171, src/reorder.h:785-980) are computed
 * exactly here: input tile counts by the device tile conversion, C tile
 * counts by the step-1 tile-level symbolic count.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/tilespgemm_reorder.h"
#include "../../include/tsg.h"

tsg_ctx *tsg_compat_ctx(void);

static double now_ms(void) {
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec * 1000.0 + t.tv_usec / 1000.0;
}

void axis_permutation_init(AxisPermutation *perm) {
    perm->n = 0;
    perm->old_id = NULL;
    perm->new_id = NULL;
}

void axis_permutation_destroy(AxisPermutation *perm) {
    free(perm->old_id);
    free(perm->new_id);
    axis_permutation_init(perm);
}

int axis_permutation_alloc(AxisPermutation *perm, int n) {
    axis_permutation_destroy(perm);
    perm->n = n;
    perm->old_id = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    perm->new_id = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!perm->old_id || !perm->new_id) return -1;
    return 0;
}

int axis_permutation_identity(AxisPermutation *perm, int n) {
    if (axis_permutation_alloc(perm, n) != 0) return -1;
    for (int i = 0; i < n; i++) perm->old_id[i] = perm->new_id[i] = i;
    return 0;
}

int axis_permutation_copy(AxisPermutation *dst, const AxisPermutation *src) {
    if (axis_permutation_alloc(dst, src->n) != 0) return -1;
    memcpy(dst->old_id, src->old_id, (size_t)src->n * sizeof(int));
    memcpy(dst->new_id, src->new_id, (size_t)src->n * sizeof(int));
    return 0;
}

int axis_permutation_is_identity(const AxisPermutation *perm) {
    if (!perm || !perm->old_id || !perm->new_id) return 0;
    for (int i = 0; i < perm->n; i++)
        if (perm->old_id[i] != i || perm->new_id[i] != i) return 0;
    return 1;
}

static void set_text(char *dst, size_t n, const char *src) {
    if (!n) return;
    strncpy(dst, src ? src : "", n - 1);
    dst[n - 1] = '\0';
}

void reorder_info_init(ReorderInfo *info) {
    memset(info, 0, sizeof(*info));
    axis_permutation_init(&info->row_perm);
    axis_permutation_init(&info->inner_perm);
    axis_permutation_init(&info->col_perm);
    info->risk_input_tile_ratio = 1.0;
    info->risk_c_tile_ratio = 1.0;
    set_text(info->mode, sizeof(info->mode), "unset");
    set_text(info->risk_reason, sizeof(info->risk_reason), "none");
}

void reorder_info_destroy(ReorderInfo *info) {
    axis_permutation_destroy(&info->row_perm);
    axis_permutation_destroy(&info->inner_perm);
    axis_permutation_destroy(&info->col_perm);
    reorder_info_init(info);
}

/* ------------------------------------------------------------------ device helpers */
static int upload(tsg_ctx *c, const SMatrix *s, tsg_dcsr *d) {
    tsg_hcsr h = {s->m, s->n, s->nnz, s->rowpointer, s->columnindex, s->value};
    return tsg_dcsr_upload(c, &h, d);
}

/* order (old_id) of an n x n matrix by `algo`, into perm (host, both maps) */
static int device_order(tsg_ctx *c, const SMatrix *s, const char *algo, AxisPermutation *perm) {
    const int n = s->m;
    if (axis_permutation_alloc(perm, n) != 0) return -1;
    if (!strcmp(algo, "original")) {
        for (int i = 0; i < n; i++) perm->old_id[i] = perm->new_id[i] = i;
        return 0;
    }
    int *d_old = NULL, *d_new = NULL;
    int st = tsg_dmalloc(c, (void **)&d_old, (size_t)(n + 1) * 4);
    if (!st) st = tsg_dmalloc(c, (void **)&d_new, (size_t)(n + 1) * 4);
    tsg_dcsr d;
    memset(&d, 0, sizeof(d));
    int have_d = 0;
    if (!st && strcmp(algo, "random") != 0) {
        st = upload(c, s, &d);
        have_d = st == 0;
    }
    if (!st) {
        if (!strcmp(algo, "rcm")) {
            st = tsg_order_rcm(c, &d, d_old, NULL);
        } else if (!strcmp(algo, "degree")) {
            st = tsg_order_degree(c, &d, s->isSymmetric, d_old);
        } else if (!strcmp(algo, "random")) {
            const char *e = getenv("TSG_RANDOM_SEED");
            st = tsg_order_random(c, n, e ? strtoull(e, NULL, 10) : 1ULL, d_old);
        } else {
            st = -20;
        }
    }
    if (!st) st = tsg_perm_invert(c, n, d_old, d_new);
    if (!st) st = tsg_d2h(c, perm->old_id, d_old, (size_t)n * 4);
    if (!st) st = tsg_d2h(c, perm->new_id, d_new, (size_t)n * 4);
    if (have_d) tsg_dcsr_free(c, &d);
    tsg_dfree(c, d_old);
    tsg_dfree(c, d_new);
    if (st == -20) return -20;
    return st ? -30 : 0;
}

/* P*A*Q^T on the device; result downloaded into out (host arrays) or kept
 * on the device in dout. */
static int device_permute(tsg_ctx *c, const SMatrix *in, const int *row_old, const int *col_new, tsg_dcsr *dout) {
    tsg_dcsr d;
    int *dr = NULL, *dc = NULL;
    int st = upload(c, in, &d);
    if (st) return st;
    st = tsg_dmalloc(c, (void **)&dr, (size_t)(in->m + 1) * 4);
    if (!st) st = tsg_dmalloc(c, (void **)&dc, (size_t)(in->n + 1) * 4);
    if (!st) st = tsg_h2d(c, dr, row_old, (size_t)in->m * 4);
    if (!st) st = tsg_h2d(c, dc, col_new, (size_t)in->n * 4);
    if (!st) st = tsg_dcsr_permute(c, &d, dr, dc, dout);
    tsg_dfree(c, dr);
    tsg_dfree(c, dc);
    tsg_dcsr_free(c, &d);
    return st;
}

/* exact tile statistics for the info block */
static long long count_tiles_dev(tsg_ctx *c, const tsg_dcsr *d) {
    tsg_dtiles t;
    if (tsg_csr2tile(c, d, &t) != TSG_OK) return -1;
    long long n = t.numtile;
    tsg_tiles_free(c, &t);
    return n;
}

static long long count_c_tiles_dev(tsg_ctx *c, const tsg_dcsr *a, const tsg_dcsr *b) {
    tsg_dtiles ta, tb;
    if (tsg_csr2tile(c, a, &ta) != TSG_OK) return -1;
    if (tsg_csr2tile(c, b, &tb) != TSG_OK) { tsg_tiles_free(c, &ta); return -1; }
    long long n = -1;
    if (tsg_spgemm_count_tiles(c, &ta, &tb, &n) != TSG_OK) n = -1;
    tsg_tiles_free(c, &ta);
    tsg_tiles_free(c, &tb);
    return n;
}

static double ratio(long long a, long long b) { return b > 0 ? (double)a / (double)b : 1.0; }

static void fill_estimates(tsg_ctx *c, ReorderProblemKind kind, const SMatrix *A, const SMatrix *B,
                           ReorderInfo *info) {
    tsg_dcsr a0, b0, a1, b1;
    if (upload(c, A, &a0)) return;
    if (upload(c, B, &b0)) { tsg_dcsr_free(c, &a0); return; }
    if (device_permute(c, A, info->row_perm.old_id, info->inner_perm.new_id, &a1)) goto out0;
    if (device_permute(c, B, info->inner_perm.old_id, info->col_perm.new_id, &b1)) { tsg_dcsr_free(c, &a1); goto out0; }
    info->input_tiles_A_before = count_tiles_dev(c, &a0);
    info->input_tiles_A_after = count_tiles_dev(c, &a1);
    if (kind == REORDER_PROBLEM_AA) {
        info->input_tiles_B_before = info->input_tiles_A_before;
        info->input_tiles_B_after = info->input_tiles_A_after;
    } else {
        info->input_tiles_B_before = count_tiles_dev(c, &b0);
        info->input_tiles_B_after = count_tiles_dev(c, &b1);
    }
    info->estimated_c_tiles_before = count_c_tiles_dev(c, &a0, &b0);
    info->estimated_c_tiles_after = count_c_tiles_dev(c, &a1, &b1);
    {
        double ra = ratio(info->input_tiles_A_after, info->input_tiles_A_before);
        double rb = ratio(info->input_tiles_B_after, info->input_tiles_B_before);
        info->risk_input_tile_ratio = ra > rb ? ra : rb;
        info->risk_c_tile_ratio = ratio(info->estimated_c_tiles_after, info->estimated_c_tiles_before);
    }
    tsg_dcsr_free(c, &a1);
    tsg_dcsr_free(c, &b1);
out0:
    tsg_dcsr_free(c, &a0);
    tsg_dcsr_free(c, &b0);
}

int build_external_reorder_plan(ReorderProblemKind kind, const char *algorithm, const char *matrix_path_A,
                                const char *matrix_path_B, const SMatrix *matrixA, const SMatrix *matrixB,
                                ReorderInfo *info, double *ordergen_ms) {
    (void)matrix_path_A;
    (void)matrix_path_B;
    if (!matrixA || !info || !ordergen_ms || !algorithm) return -1;
    if (kind != REORDER_PROBLEM_AA && !matrixB) return -2;
    if (kind != REORDER_PROBLEM_AA && matrixA->n != matrixB->m) return -3;
    reorder_info_destroy(info);
    *ordergen_ms = 0.0;
    tsg_ctx *c = tsg_compat_ctx();
    const double t0 = now_ms();
    int st = 0;
    if (kind == REORDER_PROBLEM_AA) {
        if (matrixA->m != matrixA->n) return -4;
        st = device_order(c, matrixA, algorithm, &info->row_perm);
        if (st) return st;
        if (axis_permutation_copy(&info->inner_perm, &info->row_perm) ||
            axis_permutation_copy(&info->col_perm, &info->row_perm))
            return -1;
        info->shared_symmetric_csr = 1;
    } else if (kind == REORDER_PROBLEM_AAT) {
        if (matrixA->m != matrixA->n || matrixB->m != matrixB->n || matrixA->m != matrixB->m) return -5;
        st = device_order(c, matrixA, algorithm, &info->row_perm);
        if (st) return st;
        if (axis_permutation_copy(&info->inner_perm, &info->row_perm) ||
            axis_permutation_copy(&info->col_perm, &info->row_perm))
            return -1;
        info->shared_symmetric_csr = 0;
    } else if (kind == REORDER_PROBLEM_AB) {
        if (matrixA->m != matrixA->n) return -6;
        if (matrixB->m != matrixB->n) return -7;
        if (matrixA->n != matrixB->m) return -8;
        st = device_order(c, matrixA, algorithm, &info->row_perm);
        if (st) return st;
        if (axis_permutation_copy(&info->inner_perm, &info->row_perm)) return -1;
        st = device_order(c, matrixB, algorithm, &info->col_perm);
        if (st) return st;
        info->shared_symmetric_csr = 0;
    } else {
        return -9;
    }
    *ordergen_ms = now_ms() - t0;
    info->enabled = 1;
    info->n = matrixA->n;
    info->guard_degraded_to_identity = 0;
    info->structural_conservative = 0;
    set_text(info->mode, sizeof(info->mode), algorithm);
    set_text(info->risk_reason, sizeof(info->risk_reason), "device_order");
    fill_estimates(c, kind, matrixA, kind == REORDER_PROBLEM_AA ? matrixA : matrixB, info);
    info->reorder_time_ms = now_ms() - t0;
    return 0;
}

int build_original_reorder_plan(ReorderProblemKind kind, const SMatrix *matrixA, const SMatrix *matrixB,
                                ReorderInfo *info) {
    if (!matrixA || !info) return -1;
    if (kind != REORDER_PROBLEM_AA && !matrixB) return -2;
    reorder_info_destroy(info);
    if (axis_permutation_identity(&info->row_perm, matrixA->m) ||
        axis_permutation_identity(&info->inner_perm, matrixA->n) ||
        axis_permutation_identity(&info->col_perm, kind == REORDER_PROBLEM_AA ? matrixA->n : matrixB->n))
        return -1;
    info->enabled = 1;
    info->n = matrixA->n;
    info->shared_symmetric_csr = kind == REORDER_PROBLEM_AA;
    info->guard_degraded_to_identity = 1;
    set_text(info->mode, sizeof(info->mode), "original");
    set_text(info->risk_reason, sizeof(info->risk_reason), "original_order");
    fill_estimates(tsg_compat_ctx(), kind, matrixA, kind == REORDER_PROBLEM_AA ? matrixA : matrixB, info);
    return 0;
}

int build_reorder_plan(const ReorderProblem *problem, ReorderInfo *info) {
    if (!problem || !info || !problem->matrixA) return -1;
    if (problem->kind != REORDER_PROBLEM_AA && !problem->matrixB) return -3;
    if (problem->kind > REORDER_PROBLEM_AB) return -2;
    double ms = 0.0;
    return build_external_reorder_plan(problem->kind, "rcm", NULL, NULL, problem->matrixA, problem->matrixB, info,
                                       &ms);
}

int apply_bipermutation_csr(const SMatrix *input, const AxisPermutation *row_perm, const AxisPermutation *col_perm,
                            SMatrix *output) {
    if (!row_perm || !col_perm || row_perm->n != input->m || col_perm->n != input->n) return -1;
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dcsr d;
    if (device_permute(c, input, row_perm->old_id, col_perm->new_id, &d)) return -2;
    tsg_hcsr h;
    int st = tsg_dcsr_download(c, &d, &h);
    tsg_dcsr_free(c, &d);
    if (st) return -2;
    output->m = input->m;
    output->n = input->n;
    output->nnz = input->nnz;
    output->isSymmetric = 0;
    output->rowpointer = h.rowptr;
    output->columnindex = h.colidx;
    output->value = h.val;
    return 0;
}

int restore_bipermutation_csr(SMatrix *matrix, const AxisPermutation *row_perm, const AxisPermutation *col_perm) {
    if (!row_perm || !col_perm) return 0;
    if (row_perm->n != matrix->m || col_perm->n != matrix->n) return -1;
    tsg_ctx *c = tsg_compat_ctx();
    tsg_dcsr d;
    /* restored row R = current row new_id[R]; current column c -> old_id[c] */
    if (device_permute(c, matrix, row_perm->new_id, col_perm->old_id, &d)) return -1;
    tsg_hcsr h;
    int st = tsg_dcsr_download(c, &d, &h);
    tsg_dcsr_free(c, &d);
    if (st) return -1;
    free(matrix->rowpointer);
    free(matrix->columnindex);
    free(matrix->value);
    matrix->rowpointer = h.rowptr;
    matrix->columnindex = h.colidx;
    matrix->value = h.val;
    return 0;
}
