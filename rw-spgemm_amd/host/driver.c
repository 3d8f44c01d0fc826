/*
 * driver.c -- the reference's command-line driver (src/main.cu:15-638) as a
 * library entry point, tsg_driver_main(argc, argv); the tsg_driver
 * executable (host/main.c) only calls it.
 *
 *   tsg_driver -d <dev> -aat {0|1} <A> [-b <B>] [-reorder {0|1}]
 *              [-reorder_algo tasa|rcm|degree|random|original|<adapter>]
 *              [-keepvalues] [-check {0|1}] [-o <C.bin>]
 *   <A>, <B> = matrix.mtx | gen:5pt:NX:NY | gen:27p:N | gen:27d:N
 *            | gen:er:N:ROWMEAN:SEED | gen:banded:N:HALFBW | gen:rmat:SCALE:EF:SEED
 *
 * The flow is the reference's, step for step:
 *  - load A (and B with -b) with mmio_allinone (src/main.cu:150-189);
 *  - A^2 needs a square A, A*B refuses -aat 1 and needs colA == rowB;
 *  - values overwritten with i % 10 (src/main.cu:191-197; -keepvalues
 *    keeps them);
 *  - B = A (A^2), B = A^T (AA^T, symmetric inputs refused,
 *    src/main.cu:198-241) or the -b matrix;
 *  - -reorder 1: a plan from build_reorder_plan ("tasa"),
 *    build_original_reorder_plan ("original") or build_external_reorder_plan
 *    (anything else), applied to A and B with apply_bipermutation_csr, the
 *    AA problem sharing one permuted CSR (src/main.cu:245-360);
 *  - nnzCub, csr2tile_row_major(A), csr2tile_col_major(B), tilespgemm,
 *    tile2csr(C), and restore_bipermutation_csr(C) when a permutation was
 *    applied (src/main.cu:391-581).
 * Two things are added.  Each loaded matrix gets its rows sorted and its
 * duplicates summed after the value override (the tile conversion needs
 * sorted rows; for files stored column-major this changes nothing).  The
 * result check (src/main.cu:546-602 compares with cuSPARSE) compares C
 * with a second device product of the original, unpermuted operands by
 * the tile-parallel kernels, an independent path: structure exactly, each
 * value within 1e-12 of the sum of |a_ik * b_kj| (the permutation changes
 * the summation order).  -o writes C as a binary CSR: "TSGCSR1\0", int64
 * m, n, nnz, int32 rowptr[m+1], int32 col[nnz], float64 val[nnz].
 * Returns 0, or 1 when an input cannot be loaded, 3 when the check fails.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/tilespgemm.h"
#include "../../include/tilespgemm_reorder.h"
#include "../../include/tsg.h"

int tsg_compat_set_device(int device);
tsg_ctx *tsg_compat_ctx(void);

static double now_ms(void) {
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec * 1000.0 + t.tv_usec / 1000.0;
}

static const char *kind_name(ReorderProblemKind k) {
    return k == REORDER_PROBLEM_AA ? "AA" : k == REORDER_PROBLEM_AAT ? "AAT" : k == REORDER_PROBLEM_AB ? "AB" : "unknown";
}

/* Input into an SMatrix's CSR fields: a .mtx through mmio_allinone, or a
 * generator spec. */
static int load_input(const char *spec, SMatrix *s) {
    tsg_hcsr h;
    memset(&h, 0, sizeof(h));
    int st = TSG_ERR_ARG;
    s->isSymmetric = 0;
    if (strncmp(spec, "gen:", 4) != 0) {
        int m = 0, n = 0, nnz = 0, sym = 0;
        st = mmio_allinone(&m, &n, &nnz, &sym, &h.rowptr, &h.colidx, &h.val, (char *)spec);
        if (st != 0) return st;
        h.m = m; h.n = n; h.nnz = nnz;
        s->isSymmetric = sym;
    } else {
        char kind[32] = {0};
        long long a1 = 0, a2 = 0, a3 = 0;
        double d = 0;
        if (sscanf(spec, "gen:%31[^:]", kind) != 1) return TSG_ERR_ARG;
        const char *rest = spec + 4 + strlen(kind);
        if (!strcmp(kind, "5pt") && sscanf(rest, ":%lld:%lld", &a1, &a2) == 2)
            st = tsg_gen_stencil5((int)a1, (int)a2, &h);
        else if ((!strcmp(kind, "27p") || !strcmp(kind, "27d")) && sscanf(rest, ":%lld", &a1) == 1)
            st = tsg_gen_stencil27((int)a1, (int)a1, (int)a1, kind[2] == 'p', TSG_VAL_UNIFORM_PM1, 1, &h);
        else if (!strcmp(kind, "er") && sscanf(rest, ":%lld:%lf:%lld", &a1, &d, &a3) == 3)
            st = tsg_gen_er((int)a1, d, TSG_VAL_UNIFORM_PM1, (unsigned long long)a3, &h);
        else if (!strcmp(kind, "banded") && sscanf(rest, ":%lld:%lld", &a1, &a2) == 2)
            st = tsg_gen_banded((int)a1, (int)a2, TSG_VAL_UNIFORM_PM1, 1, &h);
        else if (!strcmp(kind, "rmat") && sscanf(rest, ":%lld:%lld:%lld", &a1, &a2, &a3) == 3)
            st = tsg_gen_rmat((int)a1, (int)a2, 0.57, 0.19, 0.19, 1, TSG_VAL_UNIFORM_01, (unsigned long long)a3, &h);
        if (st != TSG_OK) return st;
    }
    s->m = h.m;
    s->n = h.n;
    s->nnz = h.nnz;
    s->rowpointer = h.rowptr;
    s->columnindex = h.colidx;
    s->value = h.val;
    return 0;
}

/* rows sorted, duplicates summed (nnz may shrink) */
static void canonicalize(SMatrix *s) {
    tsg_hcsr h = {s->m, s->n, s->nnz, s->rowpointer, s->columnindex, s->value};
    if (tsg_hcsr_canonicalize(&h) == TSG_OK) s->nnz = h.nnz;
}

static void free_csr(SMatrix *s) {
    free(s->rowpointer);
    free(s->columnindex);
    free(s->value);
    s->rowpointer = NULL;
    s->columnindex = NULL;
    s->value = NULL;
}

static int write_csr(const char *path, const SMatrix *c) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const int64_t hdr[3] = {c->m, c->n, c->nnz};
    int ok = fwrite("TSGCSR1", 1, 8, f) == 8 && fwrite(hdr, 8, 3, f) == 3 &&
             fwrite(c->rowpointer, 4, (size_t)c->m + 1, f) == (size_t)c->m + 1 &&
             fwrite(c->columnindex, 4, (size_t)c->nnz, f) == (size_t)c->nnz &&
             fwrite(c->value, 8, (size_t)c->nnz, f) == (size_t)c->nnz;
    return (fclose(f) == 0 && ok) ? 0 : -1;
}

/* C (host CSR, original order) against the tile-parallel device product of
 * the original operands.  Returns 0 when it matches. */
static int check_result(const SMatrix *A, const SMatrix *B, const SMatrix *C) {
    tsg_ctx *c = tsg_compat_ctx();
    tsg_hcsr ha = {A->m, A->n, A->nnz, A->rowpointer, A->columnindex, A->value};
    tsg_hcsr hb = {B->m, B->n, B->nnz, B->rowpointer, B->columnindex, B->value};
    double *absa = (double *)malloc(sizeof(double) * (size_t)(A->nnz ? A->nnz : 1));
    double *absb = (double *)malloc(sizeof(double) * (size_t)(B->nnz ? B->nnz : 1));
    if (!absa || !absb) { free(absa); free(absb); return -1; }
    for (int i = 0; i < A->nnz; i++) absa[i] = fabs(A->value[i]);
    for (int i = 0; i < B->nnz; i++) absb[i] = fabs(B->value[i]);
    tsg_hcsr haa = ha, hbb = hb;
    haa.val = absa;
    hbb.val = absb;
    tsg_tune tune;
    tsg_tune_default(&tune);
    tune.force_tile_parallel = 1;
    tune.element_mode = 1;
    tsg_hcsr ref[2];
    memset(ref, 0, sizeof(ref));
    int st = 0;
    for (int pass = 0; pass < 2 && !st; pass++) {
        tsg_dcsr da, db, dc;
        tsg_dtiles ta, tb, tc;
        st = tsg_dcsr_upload(c, pass ? &haa : &ha, &da);
        if (st) break;
        st = tsg_dcsr_upload(c, pass ? &hbb : &hb, &db);
        if (!st) st = tsg_csr2tile(c, &da, &ta);
        if (!st) st = tsg_csr2tile(c, &db, &tb);
        if (!st) st = tsg_tiles_build_csc(c, &tb);
        if (!st) st = tsg_spgemm(c, &ta, &tb, &tune, &tc, NULL);
        if (!st) st = tsg_tile2csr(c, &tc, &dc);
        if (!st) st = tsg_dcsr_download(c, &dc, &ref[pass]);
        if (!st) { tsg_dcsr_free(c, &dc); tsg_tiles_free(c, &tc); }
        tsg_tiles_free(c, &tb);
        tsg_tiles_free(c, &ta);
        tsg_dcsr_free(c, &db);
        tsg_dcsr_free(c, &da);
    }
    free(absa);
    free(absb);
    if (st) {
        printf("check: reference product failed: %s\n", tsg_last_error());
        tsg_hcsr_free(&ref[0]);
        tsg_hcsr_free(&ref[1]);
        return -1;
    }
    int bad = ref[0].m != C->m || ref[0].nnz != C->nnz;
    for (int i = 0; !bad && i <= C->m; i++) bad = ref[0].rowptr[i] != C->rowpointer[i];
    long long vbad = 0;
    double maxerr = 0.0;
    for (long long i = 0; !bad && i < C->nnz; i++) {
        if (ref[0].colidx[i] != C->columnindex[i]) { bad = 1; break; }
        const double e = fabs(ref[0].val[i] - C->value[i]);
        const double scale = ref[1].val[i];
        if (e > 1e-12 * scale) vbad++;
        if (scale > 0 && e / scale > maxerr) maxerr = e / scale;
    }
    if (bad) printf("check: C structure differs from the reference product\n");
    else printf("check: structure identical, max |dc|/sum|ab| = %.3e, %lld values off\n", maxerr, vbad);
    printf("check %s\n", (bad || vbad) ? "FAILED" : "passed");
    tsg_hcsr_free(&ref[0]);
    tsg_hcsr_free(&ref[1]);
    return (bad || vbad) ? 1 : 0;
}

static void usage(void) {
    printf("Run the code by './tsg_driver -d 0 -aat 0 matrixA.mtx [-b matrixB.mtx] [-reorder 1] "
           "[-reorder_algo tasa|degree|rcm|...]'.\n");
}

int tsg_driver_main(int argc, char **argv) {
    if (argc < 6) {
        usage();
        return 0;
    }
    int argi = 1, device = 0, aat = 0, reorder = 0, keep = 0, do_check = 1;
    const char *reorder_algo = "tasa";
    const char *filename = NULL, *filename_b = NULL, *out_path = NULL;
    if (strcmp(argv[argi++], "-d") != 0) return 0;
    device = atoi(argv[argi++]);
    if (strcmp(argv[argi++], "-aat") != 0) return 0;
    aat = atoi(argv[argi++]);
    while (argi < argc) {
        const char *a = argv[argi++];
        const int has = argi < argc;
        if (!strcmp(a, "-reorder") || !strcmp(a, "-b") || !strcmp(a, "-reorder_algo") || !strcmp(a, "-check") ||
            !strcmp(a, "-o")) {
            if (!has) {
                printf("Missing value for %s.\n", a);
                return 0;
            }
            const char *v = argv[argi++];
            if (!strcmp(a, "-reorder")) reorder = atoi(v);
            else if (!strcmp(a, "-b")) filename_b = v;
            else if (!strcmp(a, "-reorder_algo")) reorder_algo = v;
            else if (!strcmp(a, "-check")) do_check = atoi(v);
            else out_path = v;
        } else if (!strcmp(a, "-keepvalues")) {
            keep = 1;
        } else {
            filename = a;
        }
    }
    if (!filename) {
        usage();
        return 0;
    }
    tsg_compat_set_device(device);
    printf("device_id = %i\n", device);

    SMatrix A0, B0, Ar, Br, C;
    memset(&A0, 0, sizeof(A0));
    memset(&B0, 0, sizeof(B0));
    memset(&Ar, 0, sizeof(Ar));
    memset(&Br, 0, sizeof(Br));
    memset(&C, 0, sizeof(C));
    const int ab_mode = filename_b != NULL;
    int b_owns_csr = 0;
    ReorderProblemKind kind = REORDER_PROBLEM_AA;

    printf("MAT A: -------------- %s --------------\n", filename);
    if (ab_mode) printf("MAT B: -------------- %s --------------\n", filename_b);
    double t0 = now_ms();
    int st = load_input(filename, &A0);
    if (st != 0) {
        printf("cannot load %s (%d)\n", filename, st);
        return 1;
    }
    printf("input matrix A: ( %i, %i ) nnz = %i\n loadfile time    = %4.5f sec\n", A0.m, A0.n, A0.nnz,
           (now_ms() - t0) / 1000.0);
    int rc = 0;
    if (!aat && !ab_mode && A0.m != A0.n) {
        printf("matrix squaring must have rowA == colA. Exit.\n");
        goto done;
    }
    if (ab_mode) {
        if (aat) {
            printf("explicit A*B mode does not support -aat 1. Exit.\n");
            goto done;
        }
        kind = REORDER_PROBLEM_AB;
        t0 = now_ms();
        st = load_input(filename_b, &B0);
        if (st != 0) {
            printf("cannot load %s (%d)\n", filename_b, st);
            rc = 1;
            goto done;
        }
        b_owns_csr = 1;
        printf("input matrix B: ( %i, %i ) nnz = %i\n loadfile time    = %4.5f sec\n", B0.m, B0.n, B0.nnz,
               (now_ms() - t0) / 1000.0);
        if (A0.n != B0.m) {
            printf("A*B requires colA == rowB. Exit.\n");
            goto done;
        }
    }
    printf("the tilesize = %d\n", TSG_BLOCK_SIZE);
    if (!keep) {
        for (int i = 0; i < A0.nnz; i++) A0.value[i] = i % 10;
        if (ab_mode)
            for (int i = 0; i < B0.nnz; i++) B0.value[i] = i % 10;
    }
    canonicalize(&A0);
    if (ab_mode) canonicalize(&B0);
    if (!ab_mode && aat) {
        if (A0.m == A0.n && A0.isSymmetric) {
            printf("matrix AAT does not do symmetric matrix. Exit.\n");
            goto done;
        }
        kind = REORDER_PROBLEM_AAT;
        B0.m = A0.n;
        B0.n = A0.m;
        B0.nnz = A0.nnz;
        B0.isSymmetric = 0;
        B0.rowpointer = (int *)malloc(((size_t)A0.n + 1) * sizeof(int));
        B0.columnindex = (int *)malloc((size_t)A0.nnz * sizeof(int) + 4);
        B0.value = (double *)malloc((size_t)A0.nnz * sizeof(double) + 8);
        b_owns_csr = 1;
        matrix_transposition(A0.m, A0.n, A0.nnz, A0.rowpointer, A0.columnindex, A0.value, B0.columnindex,
                             B0.rowpointer, B0.value);
    } else if (!ab_mode) {
        B0 = A0; /* B aliases A's CSR */
    }

    ReorderInfo info;
    reorder_info_init(&info);
    SMatrix *A = &A0, *B = &B0;
    SMatrix Bt; /* tile holder of B when B aliases A's CSR */
    int materialized = 0, shared_csr = 0;
    double ordergen_ms = 0.0;
    if (reorder) {
        printf("reorder enabled\n");
        printf("reorder problem = %s\n", kind_name(kind));
        printf("reorder algorithm = %s\n", reorder_algo);
        printf("reorder strategy = %s\n", !strcmp(reorder_algo, "tasa") ? "TASA-Fast-Unified" : "External-Order");
        t0 = now_ms();
        SMatrix *rb = kind == REORDER_PROBLEM_AA ? &A0 : &B0;
        if (!strcmp(reorder_algo, "tasa")) {
            ReorderProblem p = {kind, &A0, rb};
            st = build_reorder_plan(&p, &info);
            if (st) { printf("reorder failed, status = %d. Exit.\n", st); goto done_info; }
        } else if (!strcmp(reorder_algo, "original")) {
            st = build_original_reorder_plan(kind, &A0, rb, &info);
            if (st) { printf("original reorder plan failed, status = %d. Exit.\n", st); goto done_info; }
        } else {
            st = build_external_reorder_plan(kind, reorder_algo, filename, filename_b, &A0, rb, &info, &ordergen_ms);
            if (st) { printf("external reorder failed, status = %d. Exit.\n", st); goto done_info; }
        }
        if (!info.guard_degraded_to_identity) {
            const double tp = now_ms();
            st = apply_bipermutation_csr(&A0, &info.row_perm, &info.inner_perm, &Ar);
            if (st) { printf("reorder apply A failed, status = %d. Exit.\n", st); goto done_info; }
            if (kind == REORDER_PROBLEM_AA && info.shared_symmetric_csr) {
                Br = Ar;
                shared_csr = 1;
            } else {
                st = apply_bipermutation_csr(kind == REORDER_PROBLEM_AA ? &A0 : &B0, &info.inner_perm,
                                             &info.col_perm, &Br);
                if (st) { printf("reorder apply B failed, status = %d. Exit.\n", st); goto done_info; }
            }
            info.permute_time_ms = now_ms() - tp;
            A = &Ar;
            B = &Br;
            materialized = 1;
        } else {
            shared_csr = kind == REORDER_PROBLEM_AA;
        }
        info.reorder_time_ms = now_ms() - t0;
        printf("reorder time = %.2f ms\n", info.reorder_time_ms);
        printf("reorder mode = %s\n", info.mode);
        printf("reorder ordergen time = %.2f ms\n", ordergen_ms);
        printf("reorder permute time = %.2f ms\n", info.permute_time_ms);
        printf("reorder preprocess time = %.2f ms\n", ordergen_ms + info.permute_time_ms);
        printf("reorder shared csr = %d\n", shared_csr);
        printf("risk reason = %s\n", info.risk_reason);
        printf("risk input tile ratio = %.4f\n", info.risk_input_tile_ratio);
        printf("risk C tile ratio = %.4f\n", info.risk_c_tile_ratio);
        if (info.guard_degraded_to_identity) printf("reorder uses conservative identity\n");
        printf("input tiles A before reorder = %lld\n", info.input_tiles_A_before);
        printf("input tiles A after reorder = %lld\n", info.input_tiles_A_after);
        printf("input tiles B before reorder = %lld\n", info.input_tiles_B_before);
        printf("input tiles B after reorder = %lld\n", info.input_tiles_B_after);
        printf("estimated C tiles before reorder = %lld\n", info.estimated_c_tiles_before);
        printf("estimated C tiles after reorder = %lld\n", info.estimated_c_tiles_after);
    }

    {
        /* B's tile arrays live in their own SMatrix even when its CSR is A's */
        Bt = *B;
        unsigned long long nnzCub = tsg_nnzcub(A, &Bt);
        printf("SpGEMM nnzCub = %lld\n", nnzCub);
        t0 = now_ms();
        csr2tile_row_major(A);
        printf("CSR to Tile conversion uses %.2f ms\n", now_ms() - t0);
        const double tile_bytes = (A->tilem + 1) * 4.0 + A->numtile * 4.0 + (A->numtile + 1) * 4.0 + A->nnz * 9.0 +
                                  A->numtile * 16.0 * 3.0;
        printf("tile space overhead = %.2f MB\n", tile_bytes / 1024 / 1024);
        csr2tile_col_major(&Bt);
        unsigned long long nnzC = 0;
        double cr = 0, t = 0, gf = 0, s1 = 0, s2 = 0, s3 = 0, sm = 0;
        tilespgemm(A, &Bt, &C, NULL, NULL, 0, 0.0, 0.0, nnzCub, &nnzC, &cr, &t, &gf, (char *)filename, &s1, &s2, &s3,
                   &sm);
        tile2csr(&C);
        printf("tile to CSR conversion complete! nnz(C) = %d, compression rate %.2f\n", C.nnz, cr);
        matrix_destroy(A);
        matrix_destroy(&Bt);
    }
    if (materialized) {
        t0 = now_ms();
        st = restore_bipermutation_csr(&C, &info.row_perm, &info.col_perm);
        if (st) { printf("restore reorder result failed, status = %d. Exit.\n", st); goto done_c; }
        printf("restore reorder result uses %.2f ms\n", now_ms() - t0);
    }
    if (do_check) {
        printf("-------------------------------check----------------------------------------\n");
        if (check_result(&A0, kind == REORDER_PROBLEM_AA ? &A0 : &B0, &C) != 0) rc = 3;
    }
    if (out_path && write_csr(out_path, &C) != 0) {
        printf("cannot write %s\n", out_path);
        rc = 1;
    }
done_c:
    free(C.tile_ptr); free(C.tile_columnidx); free(C.tile_nnz); free(C.tile_csr_Value);
    free(C.tile_csr_Col); free(C.tile_csr_Ptr);
    free_csr(&C);
done_info:
    if (materialized) {
        free_csr(&Ar);
        if (!shared_csr) free_csr(&Br);
    }
    reorder_info_destroy(&info);
done:
    free_csr(&A0);
    if (b_owns_csr) free_csr(&B0);
    return rc;
}
