/*
 * main.c -- tsg_driver, the command-line driver of the reference
 * (src/main.cu:15-638) on the MI355X library:
 *
 *   tsg_driver -d <dev> -aat {0|1} <input> [-keepvalues]
 *   <input> = matrix.mtx | gen:5pt:NX:NY | gen:27p:N | gen:27d:N
 *           | gen:er:N:ROWMEAN:SEED | gen:banded:N:HALFBW | gen:rmat:SCALE:EF:SEED
 *
 * Same flow as the reference: load, overwrite values with i % 10
 * (src/main.cu:191-192; -keepvalues skips it), B = A (A^2) or B = A^T
 * (src/main.cu:198-229, symmetric inputs refused for AA^T), nnzCub,
 * csr2tile_row_major(A), csr2tile_col_major(B), tilespgemm, tile2csr(C).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/tilespgemm.h"
#include "../../include/tsg.h"

int tsg_compat_set_device(int device);
int tsg_read_mtx(const char *path, tsg_hcsr *out, int *is_symmetric);

static double now_ms(void) {
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec * 1000.0 + t.tv_usec / 1000.0;
}

static int load_input(const char *spec, tsg_hcsr *h, int *sym) {
    *sym = 0;
    if (strncmp(spec, "gen:", 4) != 0) return tsg_read_mtx(spec, h, sym);
    char kind[32] = {0};
    long long a1 = 0, a2 = 0, a3 = 0;
    double d = 0;
    if (sscanf(spec, "gen:%31[^:]", kind) != 1) return TSG_ERR_ARG;
    const char *rest = spec + 4 + strlen(kind);
    if (!strcmp(kind, "5pt")) {
        if (sscanf(rest, ":%lld:%lld", &a1, &a2) != 2) return TSG_ERR_ARG;
        return tsg_gen_stencil5((int)a1, (int)a2, h);
    }
    if (!strcmp(kind, "27p") || !strcmp(kind, "27d")) {
        if (sscanf(rest, ":%lld", &a1) != 1) return TSG_ERR_ARG;
        return tsg_gen_stencil27((int)a1, (int)a1, (int)a1, kind[2] == 'p', TSG_VAL_UNIFORM_PM1, 1, h);
    }
    if (!strcmp(kind, "er")) {
        if (sscanf(rest, ":%lld:%lf:%lld", &a1, &d, &a3) != 3) return TSG_ERR_ARG;
        return tsg_gen_er((int)a1, d, TSG_VAL_UNIFORM_PM1, (unsigned long long)a3, h);
    }
    if (!strcmp(kind, "banded")) {
        if (sscanf(rest, ":%lld:%lld", &a1, &a2) != 2) return TSG_ERR_ARG;
        return tsg_gen_banded((int)a1, (int)a2, TSG_VAL_UNIFORM_PM1, 1, h);
    }
    if (!strcmp(kind, "rmat")) {
        if (sscanf(rest, ":%lld:%lld:%lld", &a1, &a2, &a3) != 3) return TSG_ERR_ARG;
        return tsg_gen_rmat((int)a1, (int)a2, 0.57, 0.19, 0.19, 1, TSG_VAL_UNIFORM_01, (unsigned long long)a3, h);
    }
    return TSG_ERR_ARG;
}

int main(int argc, char **argv) {
    if (argc < 6) {
        printf("Run the code by './tsg_driver -d 0 -aat 0 <matrix.mtx|gen:...> [-keepvalues]'.\n");
        return 0;
    }
    int argi = 1, device = 0, aat = 0, keep = 0;
    const char *input = NULL;
    if (strcmp(argv[argi++], "-d") != 0) return 0;
    device = atoi(argv[argi++]);
    if (strcmp(argv[argi++], "-aat") != 0) return 0;
    aat = atoi(argv[argi++]);
    for (; argi < argc; argi++) {
        if (!strcmp(argv[argi], "-keepvalues")) keep = 1;
        else input = argv[argi];
    }
    if (!input) {
        printf("missing input\n");
        return 0;
    }
    tsg_compat_set_device(device);
    printf("device_id = %i\n", device);
    printf("MAT A: -------------- %s --------------\n", input);

    tsg_hcsr h;
    int sym = 0;
    double t0 = now_ms();
    int st = load_input(input, &h, &sym);
    if (st != TSG_OK) {
        printf("cannot load %s (%d)\n", input, st);
        return 1;
    }
    printf("input matrix A: ( %i, %i ) nnz = %i\n loadfile time    = %4.5f sec\n", h.m, h.n, h.nnz,
           (now_ms() - t0) / 1000.0);
    SMatrix A, B, C;
    memset(&A, 0, sizeof(A));
    memset(&B, 0, sizeof(B));
    memset(&C, 0, sizeof(C));
    A.m = h.m; A.n = h.n; A.nnz = h.nnz; A.isSymmetric = sym;
    A.rowpointer = h.rowptr; A.columnindex = h.colidx; A.value = h.val;
    if (!aat && A.m != A.n) {
        printf("matrix squaring must have rowA == colA. Exit.\n");
        return 0;
    }
    printf("the tilesize = %d\n", TSG_BLOCK_SIZE);
    if (!keep)
        for (int i = 0; i < A.nnz; i++) A.value[i] = i % 10;
    int own_b = 0;
    if (aat) {
        if (A.m == A.n && A.isSymmetric) {
            printf("matrix AAT does not do symmetric matrix. Exit.\n");
            return 0;
        }
        B.m = A.n; B.n = A.m; B.nnz = A.nnz; B.isSymmetric = 0;
        B.rowpointer = (int *)malloc(((size_t)A.n + 1) * sizeof(int));
        B.columnindex = (int *)malloc((size_t)A.nnz * sizeof(int) + 4);
        B.value = (double *)malloc((size_t)A.nnz * sizeof(double) + 8);
        matrix_transposition(A.m, A.n, A.nnz, A.rowpointer, A.columnindex, A.value, B.columnindex, B.rowpointer,
                             B.value);
        own_b = 1;
    } else {
        B.m = A.m; B.n = A.n; B.nnz = A.nnz; B.isSymmetric = A.isSymmetric;
        B.rowpointer = A.rowpointer; B.columnindex = A.columnindex; B.value = A.value;
    }
    unsigned long long nnzCub = tsg_nnzcub(&A, &B);
    printf("SpGEMM nnzCub = %lld\n", nnzCub);
    t0 = now_ms();
    csr2tile_row_major(&A);
    printf("CSR to Tile conversion uses %.2f ms\n", now_ms() - t0);
    double tile_bytes = (A.tilem + 1) * 4.0 + A.numtile * 4.0 + (A.numtile + 1) * 4.0 + A.nnz * 9.0 +
                        A.numtile * 16.0 * 3.0;
    printf("tile space overhead = %.2f MB\n", tile_bytes / 1024 / 1024);
    csr2tile_col_major(&B);

    unsigned long long nnzC = 0;
    double cr = 0, t = 0, gf = 0, s1 = 0, s2 = 0, s3 = 0, sm = 0;
    tilespgemm(&A, &B, &C, NULL, NULL, 0, 0.0, 0.0, nnzCub, &nnzC, &cr, &t, &gf, (char *)input, &s1, &s2, &s3, &sm);
    tile2csr(&C);
    printf("tile to CSR conversion complete! nnz(C) = %d, compression rate %.2f\n", C.nnz, cr);

    matrix_destroy(&A);
    matrix_destroy(&B);
    free(C.tile_ptr); free(C.tile_columnidx); free(C.tile_nnz); free(C.tile_csr_Value);
    free(C.tile_csr_Col); free(C.tile_csr_Ptr);
    free(C.rowpointer); free(C.columnindex); free(C.value);
    if (own_b) { free(B.rowpointer); free(B.columnindex); free(B.value); }
    tsg_hcsr_free(&h);
    return 0;
}
