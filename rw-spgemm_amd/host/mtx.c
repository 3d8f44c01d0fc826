/*
 * mtx.c -- Matrix Market input.
 *
 * mmio_allinone() keeps the reference loader's semantics
 * (src/mmio_highlevel.h:593-759) entry for entry: coordinate files with
 * real / integer / pattern (1.0) / complex (real part) fields; a
 * "symmetric" or "hermitian" header expands every off-diagonal entry to
 * both triangles, the mirror copy carrying the same value; "skew-symmetric"
 * is not expanded (the reference tests only the 'S' and 'H' codes);
 * duplicates are kept; each row holds its entries in file order, an
 * off-diagonal entry of a symmetric file landing in row r and in row c at
 * the moment it is read.  Row pointers count per row, then scan
 * (src/mmio_highlevel.h:680-714).  Status codes are the reference's: -1
 * cannot open, -2 bad banner, -4 bad size line; this loader adds -5 for an
 * entry outside the matrix or a file that ends early (undefined behaviour
 * in the reference).
 *
 * tsg_read_mtx() is the library's canonical form of the same file: the
 * CSR mmio_allinone() returns with each row stably sorted by column and
 * duplicates summed in file order (tsg_hcsr_canonicalize).  For files
 * stored column-major, the SuiteSparse convention, the reference CSR is
 * already row-sorted and duplicate-free, so both agree entry for entry.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tilespgemm.h"
#include "../../include/tsg.h"

/* whole file in memory, NUL-terminated */
static char *slurp(const char *path, size_t *len) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 20, n = 0;
    char *buf = (char *)malloc(cap + 1);
    while (buf) {
        size_t got = fread(buf + n, 1, cap - n, f);
        n += got;
        if (n < cap) break;
        cap *= 2;
        char *nb = (char *)realloc(buf, cap + 1);
        if (!nb) { free(buf); buf = NULL; }
        else buf = nb;
    }
    fclose(f);
    if (!buf) return NULL;
    buf[n] = '\0';
    *len = n;
    return buf;
}

/* next line [p, end-of-line); advances *pp past the newline */
static char *next_line(char **pp) {
    char *p = *pp;
    if (!*p) return NULL;
    char *e = strchr(p, '\n');
    if (e) { *e = '\0'; *pp = e + 1; }
    else *pp = p + strlen(p);
    return p;
}

static int lower_tok(const char **s, char *out, size_t cap) {
    const char *p = *s;
    while (*p == ' ' || *p == '\t' || *p == '\r') p++;
    size_t n = 0;
    while (*p && !isspace((unsigned char)*p)) {
        if (n + 1 < cap) out[n++] = (char)tolower((unsigned char)*p);
        p++;
    }
    out[n] = '\0';
    *s = p;
    return (int)n;
}

int mmio_allinone(int *m, int *n, MAT_PTR_TYPE *nnz, int *isSymmetric, MAT_PTR_TYPE **csrRowPtr, int **csrColIdx,
                  MAT_VAL_TYPE **csrVal, char *filename) {
    size_t len = 0;
    char *buf = slurp(filename, &len);
    if (!buf) return -1;
    char *p = buf;
    char *line = next_line(&p);
    char head[64], obj[64], fmt[64], field[64], sym[64];
    const char *s = line ? line : "";
    lower_tok(&s, head, sizeof(head));
    lower_tok(&s, obj, sizeof(obj));
    lower_tok(&s, fmt, sizeof(fmt));
    lower_tok(&s, field, sizeof(field));
    lower_tok(&s, sym, sizeof(sym));
    /* banner: %%MatrixMarket matrix coordinate <field> <symmetry> */
    if (strcmp(head, "%%matrixmarket") != 0 || strcmp(obj, "matrix") != 0 || strcmp(fmt, "coordinate") != 0 ||
        (strcmp(field, "real") && strcmp(field, "integer") && strcmp(field, "pattern") && strcmp(field, "complex")) ||
        (strcmp(sym, "general") && strcmp(sym, "symmetric") && strcmp(sym, "hermitian") &&
         strcmp(sym, "skew-symmetric"))) {
        free(buf);
        return -2;
    }
    const int pattern = !strcmp(field, "pattern");
    const int complex_field = !strcmp(field, "complex");
    const int symmetric = !strcmp(sym, "symmetric") || !strcmp(sym, "hermitian");
    long long M = -1, N = -1, NZ = -1;
    while ((line = next_line(&p)) != NULL) {
        if (line[0] == '%') continue;
        const char *q = line;
        while (*q && isspace((unsigned char)*q)) q++;
        if (!*q) continue;
        if (sscanf(q, "%lld %lld %lld", &M, &N, &NZ) != 3) M = -1;
        break;
    }
    if (M < 0 || N < 0 || NZ < 0 || M > 0x7ffffffeLL || N > 0x7ffffffeLL || NZ > 0x7fffffffLL) {
        free(buf);
        return -4;
    }
    int *ri = (int *)malloc(sizeof(int) * (size_t)(NZ ? NZ : 1));
    int *ci = (int *)malloc(sizeof(int) * (size_t)(NZ ? NZ : 1));
    double *vv = (double *)malloc(sizeof(double) * (size_t)(NZ ? NZ : 1));
    long long *cnt = (long long *)calloc((size_t)M + 1, sizeof(long long));
    if (!ri || !ci || !vv || !cnt) {
        free(buf); free(ri); free(ci); free(vv); free(cnt);
        return -4;
    }
    /* entries: whitespace-separated tokens, as fscanf reads them */
    char *q = p;
    int bad = 0;
    for (long long e = 0; e < NZ && !bad; e++) {
        char *end;
        for (;;) { /* skip blanks and comment lines */
            while (*q && isspace((unsigned char)*q)) q++;
            if (*q == '%') { while (*q && *q != '\n') q++; continue; }
            break;
        }
        long long r = strtoll(q, &end, 10);
        if (end == q) { bad = 1; break; }
        q = end;
        long long c = strtoll(q, &end, 10);
        if (end == q) { bad = 1; break; }
        q = end;
        double v = 1.0;
        if (!pattern) {
            v = strtod(q, &end);
            if (end == q) { bad = 1; break; }
            q = end;
            if (complex_field) { (void)strtod(q, &end); q = end; }
        }
        if (r < 1 || r > M || c < 1 || c > N || (symmetric && c > M)) { bad = 1; break; }
        ri[e] = (int)(r - 1);
        ci[e] = (int)(c - 1);
        vv[e] = v;
        cnt[r - 1]++;
        if (symmetric && r != c) cnt[c - 1]++;
    }
    free(buf);
    long long total = 0;
    for (long long i = 0; i < M && !bad; i++) total += cnt[i];
    if (bad || total > 0x7fffffffLL) {
        free(ri); free(ci); free(vv); free(cnt);
        return bad ? -5 : -4;
    }
    int *rp = (int *)malloc(sizeof(int) * ((size_t)M + 1));
    int *col = (int *)malloc(sizeof(int) * (size_t)(total ? total : 1));
    double *val = (double *)malloc(sizeof(double) * (size_t)(total ? total : 1));
    if (!rp || !col || !val) {
        free(ri); free(ci); free(vv); free(cnt); free(rp); free(col); free(val);
        return -4;
    }
    rp[0] = 0;
    for (long long i = 0; i < M; i++) rp[i + 1] = rp[i] + (int)cnt[i];
    for (long long i = 0; i < M; i++) cnt[i] = rp[i];
    for (long long e = 0; e < NZ; e++) {
        const int r = ri[e], c = ci[e];
        long long o = cnt[r]++;
        col[o] = c;
        val[o] = vv[e];
        if (symmetric && r != c) {
            o = cnt[c]++;
            col[o] = r;
            val[o] = vv[e];
        }
    }
    free(ri); free(ci); free(vv); free(cnt);
    *m = (int)M;
    *n = (int)N;
    *nnz = (int)total;
    *isSymmetric = symmetric;
    *csrRowPtr = rp;
    *csrColIdx = col;
    *csrVal = val;
    return 0;
}

/* (column, position) pairs for a stable per-row sort */
typedef struct { int c, pos; } colpos;
static int cmp_colpos(const void *a, const void *b) {
    const colpos *x = (const colpos *)a, *y = (const colpos *)b;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return x->pos < y->pos ? -1 : x->pos > y->pos;
}

int tsg_hcsr_canonicalize(tsg_hcsr *h) {
    if (!h || h->m < 0 || (h->m && !h->rowptr)) return TSG_ERR_ARG;
    int maxlen = 0;
    for (int i = 0; i < h->m; i++)
        if (h->rowptr[i + 1] - h->rowptr[i] > maxlen) maxlen = h->rowptr[i + 1] - h->rowptr[i];
    colpos *tmp = (colpos *)malloc(sizeof(colpos) * (size_t)(maxlen ? maxlen : 1));
    double *tv = (double *)malloc(sizeof(double) * (size_t)(maxlen ? maxlen : 1));
    if (!tmp || !tv) { free(tmp); free(tv); return TSG_ERR_NOMEM; }
    int w = 0;
    for (int i = 0; i < h->m; i++) {
        const int b = h->rowptr[i], e = h->rowptr[i + 1], L = e - b;
        int sorted = 1;
        for (int k = b + 1; k < e; k++)
            if (h->colidx[k] <= h->colidx[k - 1]) { sorted = 0; break; }
        h->rowptr[i] = w;
        if (sorted) {
            for (int k = b; k < e; k++, w++) {
                h->colidx[w] = h->colidx[k];
                if (h->val) h->val[w] = h->val[k];
            }
            continue;
        }
        for (int k = 0; k < L; k++) {
            tmp[k].c = h->colidx[b + k];
            tmp[k].pos = k;
            tv[k] = h->val ? h->val[b + k] : 0.0;
        }
        qsort(tmp, (size_t)L, sizeof(colpos), cmp_colpos);
        for (int k = 0; k < L; k++) {
            const double v = tv[tmp[k].pos];
            if (k > 0 && tmp[k].c == tmp[k - 1].c) {
                if (h->val) h->val[w - 1] += v;
            } else {
                h->colidx[w] = tmp[k].c;
                if (h->val) h->val[w] = v;
                w++;
            }
        }
    }
    if (h->m >= 0 && h->rowptr) h->rowptr[h->m] = w;
    h->nnz = w;
    free(tmp);
    free(tv);
    return TSG_OK;
}

int tsg_read_mtx(const char *path, tsg_hcsr *out, int *is_symmetric) {
    memset(out, 0, sizeof(*out));
    *is_symmetric = 0;
    int m = 0, n = 0, nnz = 0, sym = 0;
    int *rp = NULL, *col = NULL;
    double *val = NULL;
    const int st = mmio_allinone(&m, &n, &nnz, &sym, &rp, &col, &val, (char *)path);
    if (st != 0) return TSG_ERR_ARG;
    out->m = m;
    out->n = n;
    out->nnz = nnz;
    out->rowptr = rp;
    out->colidx = col;
    out->val = val;
    const int c = tsg_hcsr_canonicalize(out);
    if (c != TSG_OK) { tsg_hcsr_free(out); return c; }
    *is_symmetric = sym;
    return TSG_OK;
}
