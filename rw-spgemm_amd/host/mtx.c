/*
 * mtx.c -- Matrix Market reader with the semantics of the reference's
 * mmio_allinone (src/mmio_highlevel.h:593-759): coordinate format, real /
 * integer / pattern fields (pattern -> 1.0), symmetric and skew/hermitian
 * headers expanded to both triangles, 1-based indices.  One deviation: rows
 * are sorted by column here (the reference keeps file order,
 * src/mmio_highlevel.h:734-740), because the tile conversion needs sorted
 * rows; duplicate entries are summed.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tsg.h"

typedef struct { int r, c; double v; } trip;

static int cmp_trip(const void *a, const void *b) {
    const trip *x = (const trip *)a, *y = (const trip *)b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    return x->c < y->c ? -1 : x->c > y->c;
}

int tsg_read_mtx(const char *path, tsg_hcsr *out, int *is_symmetric) {
    memset(out, 0, sizeof(*out));
    *is_symmetric = 0;
    FILE *f = fopen(path, "r");
    if (!f) return TSG_ERR_ARG;
    char line[1024];
    if (!fgets(line, sizeof(line), f)) { fclose(f); return TSG_ERR_ARG; }
    for (char *p = line; *p; p++) *p = (char)tolower((unsigned char)*p);
    if (strncmp(line, "%%matrixmarket", 14) != 0 || !strstr(line, "coordinate")) { fclose(f); return TSG_ERR_ARG; }
    const int pattern = strstr(line, "pattern") != NULL;
    const int complex_field = strstr(line, "complex") != NULL;
    const int symmetric = strstr(line, "symmetric") != NULL || strstr(line, "hermitian") != NULL;
    const int skew = strstr(line, "skew") != NULL;
    if (complex_field) { fclose(f); return TSG_ERR_ARG; }
    long long m = 0, n = 0, nz = 0;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '%') continue;
        if (sscanf(line, "%lld %lld %lld", &m, &n, &nz) == 3) break;
    }
    if (m <= 0 || n <= 0 || nz < 0 || m > 0x7fffffff || n > 0x7fffffff) { fclose(f); return TSG_ERR_ARG; }
    trip *t = (trip *)malloc(sizeof(trip) * (size_t)(symmetric ? 2 * nz + 1 : nz + 1));
    if (!t) { fclose(f); return TSG_ERR_NOMEM; }
    long long k = 0;
    for (long long e = 0; e < nz; e++) {
        long long r, c;
        double v = 1.0;
        if (!fgets(line, sizeof(line), f)) break;
        if (pattern) {
            if (sscanf(line, "%lld %lld", &r, &c) != 2) { e--; continue; }
        } else if (sscanf(line, "%lld %lld %lf", &r, &c, &v) != 3) {
            e--;
            continue;
        }
        if (r < 1 || r > m || c < 1 || c > n) { free(t); fclose(f); return TSG_ERR_ARG; }
        t[k].r = (int)(r - 1); t[k].c = (int)(c - 1); t[k].v = v; k++;
        if (symmetric && r != c) { t[k].r = (int)(c - 1); t[k].c = (int)(r - 1); t[k].v = skew ? -v : v; k++; }
    }
    fclose(f);
    qsort(t, (size_t)k, sizeof(trip), cmp_trip);
    long long w = 0;
    for (long long i = 0; i < k; i++) {
        if (w > 0 && t[w - 1].r == t[i].r && t[w - 1].c == t[i].c) t[w - 1].v += t[i].v;
        else t[w++] = t[i];
    }
    if (w > 0x7fffffffLL) { free(t); return TSG_ERR_OVERFLOW; }
    out->m = (int)m; out->n = (int)n; out->nnz = (int)w;
    out->rowptr = (int *)calloc((size_t)m + 1, sizeof(int));
    out->colidx = (int *)malloc(sizeof(int) * (size_t)(w ? w : 1));
    out->val = (double *)malloc(sizeof(double) * (size_t)(w ? w : 1));
    if (!out->rowptr || !out->colidx || !out->val) { tsg_hcsr_free(out); free(t); return TSG_ERR_NOMEM; }
    for (long long i = 0; i < w; i++) {
        out->rowptr[t[i].r + 1]++;
        out->colidx[i] = t[i].c;
        out->val[i] = t[i].v;
    }
    for (long long i = 0; i < m; i++) out->rowptr[i + 1] += out->rowptr[i];
    free(t);
    *is_symmetric = symmetric;
    return TSG_OK;
}
