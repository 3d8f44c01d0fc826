/*
 * reorder.c -- the reference's reorder API (include/tilespgemm_reorder.h)
 * in C on top of the device ordering / permutation kernels.
 *
 * Plans follow the axis-sharing rules of build_external_reorder_plan
 * (src/external_reorder.h:204-319).  The orders "rcm", "degree", "random"
 * and "original" are computed on the device; any other algorithm name goes
 * through the reference's adapter contract (reorder_algorithm/README.md):
 * $TASA_REORDER_ALGORITHM_ROOT/<algo>/order_gen <matrix.mtx> <out.order>
 * is run as a child process and its new->old order file is read back with
 * the reference's checks (src/external_reorder.h:53-171).  Every argument
 * check happens before any device work, so the status codes do not need a
 * GPU.  The before/after tile statistics the reference estimates from
 * sampled rows (src/external_reorder.h:143-171, src/reorder.h:785-980) are
 * computed exactly here when a device is present: input tile counts by the
 * device tile conversion, C tile counts by the step-1 tile-level symbolic
 * count (left at -1 without a device).
 */
#include <errno.h>
#include <spawn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <sys/wait.h>
#include <unistd.h>

#include "../../include/tilespgemm_reorder.h"
#include "../../include/tsg.h"

tsg_ctx *tsg_compat_ctx(void);
tsg_ctx *tsg_compat_ctx_try(void);
extern char **environ;

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


/* ------------------------------------------------------------------ order files */
int read_external_order_file(const char *path, int n, AxisPermutation *perm) {
    int st = axis_permutation_alloc(perm, n);
    if (st != 0) return st;
    FILE *f = path ? fopen(path, "r") : NULL;
    if (!f) {
        printf("external reorder order file open failed: %s\n", path ? path : "(null)");
        return -10;
    }
    unsigned char *seen = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!seen) { fclose(f); return -1; }
    st = 0;
    for (int new_id = 0; new_id < n && !st; new_id++) {
        long long old_ll = -1;
        if (fscanf(f, "%lld", &old_ll) != 1) {
            printf("external reorder order file too short: %s at row %d\n", path, new_id);
            st = -11;
        } else if (old_ll < 0 || old_ll >= n) {
            printf("external reorder invalid id %lld at row %d, n = %d\n", old_ll, new_id, n);
            st = -12;
        } else if (seen[old_ll]) {
            printf("external reorder duplicate id %lld in %s\n", old_ll, path);
            st = -13;
        } else {
            seen[old_ll] = 1;
            perm->old_id[new_id] = (int)old_ll;
            perm->new_id[old_ll] = new_id;
        }
    }
    if (!st) {
        long long extra;
        if (fscanf(f, "%lld", &extra) == 1) {
            printf("external reorder order file has extra entries: %s\n", path);
            st = -14;
        }
    }
    free(seen);
    fclose(f);
    return st;
}

static int is_device_algorithm(const char *a) {
    return !strcmp(a, "rcm") || !strcmp(a, "degree") || !strcmp(a, "random") || !strcmp(a, "original");
}

/* file stem with characters outside [A-Za-z0-9_-] replaced by '_' */
static void path_stem(const char *path, char *out, size_t cap) {
    const char *b = path ? strrchr(path, '/') : NULL;
    b = b ? b + 1 : (path ? path : "");
    size_t n = 0;
    for (; *b && n + 1 < cap; b++) out[n++] = *b;
    out[n] = '\0';
    char *dot = strrchr(out, '.');
    if (dot) *dot = '\0';
    for (char *p = out; *p; p++) {
        const char c = *p;
        if (!((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_' || c == '-'))
            *p = '_';
    }
    if (!out[0]) snprintf(out, cap, "matrix");
}

/* The adapter contract of reorder_algorithm/README.md: run
 * <root>/<algo>/order_gen <matrix> <order> as a child process, then read
 * the new->old order it wrote (src/external_reorder.h:100-141). */
static int run_order_generator(const char *algo, const char *matrix_path, int n, AxisPermutation *perm,
                               double *ms) {
    const char *root = getenv("TASA_REORDER_ALGORITHM_ROOT");
    const char *cache = getenv("TASA_REORDER_ORDER_DIR");
    if (!root) root = "reorder_algorithm";
    if (!cache) cache = "/tmp/tasa_reorder_orders";
    if (mkdir(cache, 0775) != 0 && errno != EEXIST) { /* the generator reports the failure */ }
    char stem[256], exe[4096], order[4096];
    path_stem(matrix_path, stem, sizeof(stem));
    snprintf(exe, sizeof(exe), "%s/%s/order_gen", root, algo);
    snprintf(order, sizeof(order), "%s/%s_%s_%lld.order", cache, algo, stem, (long long)getpid());
    const double t0 = now_ms();
    int code = 127;
    if (matrix_path) {
        char *argv[] = {exe, (char *)matrix_path, order, NULL};
        pid_t pid;
        if (posix_spawn(&pid, exe, NULL, NULL, argv, environ) == 0) {
            int wst = 0;
            while (waitpid(pid, &wst, 0) < 0 && errno == EINTR) {}
            code = WIFEXITED(wst) ? WEXITSTATUS(wst) : 128 + (WIFSIGNALED(wst) ? WTERMSIG(wst) : 0);
        }
    }
    *ms += now_ms() - t0;
    if (code != 0) {
        if (code == 42) printf("external reorder missing dependency: %s\n", algo);
        else printf("external reorder order_gen failed: algorithm = %s, exit = %d\n", algo, code);
        return -20;
    }
    const int st = read_external_order_file(order, n, perm);
    if (st != 0) return st;
    unlink(order);
    return 0;
}

/* ------------------------------------------------------------------ device helpers */
static int upload(tsg_ctx *c, const SMatrix *s, tsg_dcsr *d) {
    tsg_hcsr h = {s->m, s->n, s->nnz, s->rowpointer, s->columnindex, s->value};
    return tsg_dcsr_upload(c, &h, d);
}

/* Order (old_id) of an n x n matrix computed on the device, both maps into
 * perm.  "degree" ranks by row length: an SMatrix's CSR holds both
 * triangles of a symmetric file (mmio_allinone expands them), so its row
 * length is the stored-plus-mirrored count of order_utils.py:30-45. */
static int device_order(const SMatrix *s, const char *algo, AxisPermutation *perm, double *ms) {
    const int n = s->m;
    if (axis_permutation_alloc(perm, n) != 0) return -1;
    const double t0 = now_ms();
    if (!strcmp(algo, "original")) {
        for (int i = 0; i < n; i++) perm->old_id[i] = perm->new_id[i] = i;
        *ms += now_ms() - t0;
        return 0;
    }
    tsg_ctx *c = tsg_compat_ctx_try();
    if (!c) {
        printf("external reorder missing dependency: %s (no device)\n", algo);
        return -20;
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
            st = tsg_order_degree(c, &d, 0, d_old);
        } else {
            const char *e = getenv("TSG_RANDOM_SEED");
            st = tsg_order_random(c, n, e ? strtoull(e, NULL, 10) : 1ULL, d_old);
        }
    }
    if (!st) st = tsg_perm_invert(c, n, d_old, d_new);
    if (!st) st = tsg_d2h(c, perm->old_id, d_old, (size_t)n * 4);
    if (!st) st = tsg_d2h(c, perm->new_id, d_new, (size_t)n * 4);
    if (have_d) tsg_dcsr_free(c, &d);
    tsg_dfree(c, d_old);
    tsg_dfree(c, d_new);
    *ms += now_ms() - t0;
    if (st) printf("reorder: device %s order failed: %s\n", algo, tsg_last_error());
    return st ? -30 : 0;
}

static int make_order(const SMatrix *s, const char *algo, const char *path, AxisPermutation *perm, double *ms) {
    if (is_device_algorithm(algo)) return device_order(s, algo, perm, ms);
    return run_order_generator(algo, path, s->m, perm, ms);
}

/* P*A*Q^T on the device into dout. */
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

static void fill_estimates(ReorderProblemKind kind, const SMatrix *A, const SMatrix *B, ReorderInfo *info) {
    info->input_tiles_A_before = info->input_tiles_A_after = -1;
    info->input_tiles_B_before = info->input_tiles_B_after = -1;
    info->estimated_c_tiles_before = info->estimated_c_tiles_after = -1;
    tsg_ctx *c = tsg_compat_ctx_try();
    if (!c) return;
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
    if (!matrixA || !info || !ordergen_ms) return -1;
    if (kind != REORDER_PROBLEM_AA && !matrixB) return -2;
    if (kind != REORDER_PROBLEM_AA && matrixA->n != matrixB->m) return -3;
    reorder_info_destroy(info);
    *ordergen_ms = 0.0;
    const char *algo = algorithm ? algorithm : "";
    /* shape checks of each problem kind come before any order work */
    if (kind == REORDER_PROBLEM_AA) {
        if (matrixA->m != matrixA->n) {
            printf("external reorder AA currently requires a square matrix.\n");
            return -4;
        }
    } else if (kind == REORDER_PROBLEM_AAT) {
        if (matrixA->m != matrixA->n || matrixB->m != matrixB->n || matrixA->m != matrixB->m) {
            printf("external reorder AAT currently requires square A and A^T with matching dimensions.\n");
            return -5;
        }
    } else if (kind == REORDER_PROBLEM_AB) {
        if (matrixA->m != matrixA->n) {
            printf("external reorder AB currently requires square A so row and inner axes can share A's order.\n");
            return -6;
        }
        if (matrixB->m != matrixB->n) {
            printf("external reorder AB currently requires square B so B row and column axes can be ordered by B.\n");
            return -7;
        }
        if (matrixA->n != matrixB->m) {
            printf("external reorder AB requires colA == rowB.\n");
            return -8;
        }
    } else {
        return -9;
    }
    if (!algo[0]) {
        printf("external reorder order_gen failed: algorithm = (none), exit = 127\n");
        return -20;
    }
    int st = make_order(matrixA, algo, matrix_path_A, &info->row_perm, ordergen_ms);
    if (st) return st;
    if (axis_permutation_copy(&info->inner_perm, &info->row_perm)) return -1;
    if (kind == REORDER_PROBLEM_AB) {
        /* B's own order for the output column axis */
        st = make_order(matrixB, algo, matrix_path_B ? matrix_path_B : matrix_path_A, &info->col_perm, ordergen_ms);
        if (st) return st;
    } else if (axis_permutation_copy(&info->col_perm, &info->row_perm)) {
        return -1;
    }
    info->shared_symmetric_csr = kind == REORDER_PROBLEM_AA;
    info->enabled = 1;
    info->n = matrixA->n;
    info->guard_degraded_to_identity = 0;
    info->structural_conservative = 0;
    set_text(info->mode, sizeof(info->mode), algo);
    set_text(info->risk_reason, sizeof(info->risk_reason),
             is_device_algorithm(algo) ? "device_order" : "external_order");
    fill_estimates(kind, matrixA, kind == REORDER_PROBLEM_AA ? matrixA : matrixB, info);
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
    fill_estimates(kind, matrixA, kind == REORDER_PROBLEM_AA ? matrixA : matrixB, info);
    return 0;
}

int build_reorder_plan(const ReorderProblem *problem, ReorderInfo *info) {
    if (!problem || !info || !problem->matrixA) return -1;
    switch (problem->kind) {
    case REORDER_PROBLEM_AA:
        break;
    case REORDER_PROBLEM_AAT:
    case REORDER_PROBLEM_AB:
        if (!problem->matrixB) return -3;
        break;
    default:
        return -2;
    }
    double ms = 0.0;
    return build_external_reorder_plan(problem->kind, "rcm", NULL, NULL, problem->matrixA, problem->matrixB, info,
                                       &ms);
}

int apply_bipermutation_csr(const SMatrix *input, const AxisPermutation *row_perm, const AxisPermutation *col_perm,
                            SMatrix *output) {
    if (!input || !output || !row_perm || !col_perm || row_perm->n != input->m || col_perm->n != input->n)
        return -1;
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
    if (!matrix || row_perm->n != matrix->m || col_perm->n != matrix->n) return -1;
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
