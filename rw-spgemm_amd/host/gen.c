/*
 * gen.c -- synthetic operand generators for the TileSpGEMM workloads
 * (2D 5-point, 3D 27-point, Erdos-Renyi, banded, R-MAT).  All randomness is
 * counter-based (splitmix64 of (seed, index)), so every entry's value is a
 * pure function of the seed and its position.  Rows come out sorted with
 * unique columns, the input contract of the tile conversion (the reference
 * reads .mtx files instead, src/mmio_highlevel.h:593-759).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tsg.h"

static inline uint64_t sm64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
static inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t i) {
    return sm64(sm64(seed ^ (stream * 0xd1b54a32d192ed03ULL)) + i);
}
/* [0,1) with 53 random bits */
static inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

int tsg_hcsr_set_values(tsg_hcsr *h, int mode, unsigned long long seed) {
    if (!h || (h->nnz > 0 && !h->val)) return TSG_ERR_ARG;
    switch (mode) {
    case TSG_VAL_KEEP:
        return TSG_OK;
    case TSG_VAL_UNIFORM_PM1:
        for (int i = 0; i < h->nnz; i++) h->val[i] = 2.0 * u01(rng(seed, 1, (uint64_t)i)) - 1.0;
        return TSG_OK;
    case TSG_VAL_UNIFORM_01:
        for (int i = 0; i < h->nnz; i++) h->val[i] = 1.0 - u01(rng(seed, 2, (uint64_t)i));
        return TSG_OK;
    case TSG_VAL_MOD10:
        for (int i = 0; i < h->nnz; i++) h->val[i] = (double)(i % 10);
        return TSG_OK;
    default:
        return TSG_ERR_ARG;
    }
}

static int alloc_csr(tsg_hcsr *h, int m, int n, long long nnz) {
    memset(h, 0, sizeof(*h));
    if (nnz > 0x7fffffffLL) return TSG_ERR_OVERFLOW;
    h->m = m;
    h->n = n;
    h->nnz = (int)nnz;
    h->rowptr = (int *)malloc(((size_t)m + 1) * sizeof(int));
    h->colidx = (int *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int));
    h->val = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    if (!h->rowptr || !h->colidx || !h->val) {
        tsg_hcsr_free(h);
        return TSG_ERR_NOMEM;
    }
    return TSG_OK;
}

int tsg_gen_stencil5(int nx, int ny, tsg_hcsr *out) {
    if (nx <= 0 || ny <= 0) return TSG_ERR_ARG;
    const int n = nx * ny;
    int st = alloc_csr(out, n, n, 5LL * n);
    if (st) return st;
    int e = 0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int i = x + nx * y;
            out->rowptr[i] = e;
            if (y > 0) { out->colidx[e] = i - nx; out->val[e++] = -1.0; }
            if (x > 0) { out->colidx[e] = i - 1; out->val[e++] = -1.0; }
            out->colidx[e] = i; out->val[e++] = 4.0;
            if (x < nx - 1) { out->colidx[e] = i + 1; out->val[e++] = -1.0; }
            if (y < ny - 1) { out->colidx[e] = i + nx; out->val[e++] = -1.0; }
        }
    out->rowptr[n] = e;
    out->nnz = e;
    return TSG_OK;
}

static inline int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

int tsg_gen_stencil27(int nx, int ny, int nz, int periodic, int valmode, unsigned long long seed,
                      tsg_hcsr *out) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return TSG_ERR_ARG;
    if (periodic && (nx < 3 || ny < 3 || nz < 3)) return TSG_ERR_ARG; /* neighbours must be distinct */
    const long long nl = (long long)nx * ny * nz;
    if (nl > 0x7fffffffLL) return TSG_ERR_OVERFLOW;
    const int n = (int)nl;
    int st = alloc_csr(out, n, n, 27LL * n);
    if (st) return st;
    int e = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int i = x + nx * (y + ny * z);
                out->rowptr[i] = e;
                int cols[27], k = 0;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int xx = x + dx, yy = y + dy, zz = z + dz;
                            if (periodic) {
                                xx = wrap(xx, nx); yy = wrap(yy, ny); zz = wrap(zz, nz);
                            } else if (xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) {
                                continue;
                            }
                            cols[k++] = xx + nx * (yy + ny * zz);
                        }
                for (int a = 1; a < k; a++) { /* insertion sort (wrap breaks the order) */
                    int v = cols[a], b = a - 1;
                    while (b >= 0 && cols[b] > v) { cols[b + 1] = cols[b]; b--; }
                    cols[b + 1] = v;
                }
                for (int a = 0; a < k; a++) {
                    out->colidx[e] = cols[a];
                    out->val[e] = cols[a] == i ? 26.0 : -1.0;
                    e++;
                }
            }
    out->rowptr[n] = e;
    out->nnz = e;
    return tsg_hcsr_set_values(out, valmode, seed);
}

static int poisson(double mean, uint64_t r) {
    /* inversion, mean <= ~200 */
    double u = u01(r), p = exp(-mean), c = p;
    int k = 0;
    while (u > c && k < 100000) { k++; p *= mean / k; c += p; }
    return k;
}

int tsg_gen_er(int n, double mean_row, int valmode, unsigned long long seed, tsg_hcsr *out) {
    if (n <= 0 || mean_row < 0 || mean_row > 200) return TSG_ERR_ARG;
    int *len = (int *)malloc(sizeof(int) * (size_t)n);
    if (!len) return TSG_ERR_NOMEM;
    long long tot = 0;
    for (int i = 0; i < n; i++) {
        int k = poisson(mean_row, rng(seed, 3, (uint64_t)i));
        if (k > n) k = n;
        len[i] = k;
        tot += k;
    }
    int st = alloc_csr(out, n, n, tot);
    if (st) { free(len); return st; }
    long long e = 0;
    for (int i = 0; i < n; i++) {
        out->rowptr[i] = (int)e;
        int *row = out->colidx + e;
        int k = 0;
        uint64_t draw = 0;
        while (k < len[i]) {
            int c = (int)(rng(seed, 4 + (uint64_t)i, draw++) % (uint64_t)n);
            int a = k - 1, dup = 0;
            for (int b = 0; b < k; b++) if (row[b] == c) { dup = 1; break; }
            if (dup) continue;
            while (a >= 0 && row[a] > c) { row[a + 1] = row[a]; a--; }
            row[a + 1] = c;
            k++;
        }
        e += k;
    }
    out->rowptr[n] = (int)e;
    free(len);
    for (long long x = 0; x < e; x++) out->val[x] = 1.0;
    return tsg_hcsr_set_values(out, valmode, seed);
}

int tsg_gen_banded(int n, int halfbw, int valmode, unsigned long long seed, tsg_hcsr *out) {
    if (n <= 0 || halfbw < 0) return TSG_ERR_ARG;
    long long tot = 0;
    for (int i = 0; i < n; i++) {
        int lo = i - halfbw < 0 ? 0 : i - halfbw, hi = i + halfbw >= n ? n - 1 : i + halfbw;
        tot += hi - lo + 1;
    }
    int st = alloc_csr(out, n, n, tot);
    if (st) return st;
    int e = 0;
    for (int i = 0; i < n; i++) {
        out->rowptr[i] = e;
        int lo = i - halfbw < 0 ? 0 : i - halfbw, hi = i + halfbw >= n ? n - 1 : i + halfbw;
        for (int j = lo; j <= hi; j++) { out->colidx[e] = j; out->val[e] = 1.0; e++; }
    }
    out->rowptr[n] = e;
    return tsg_hcsr_set_values(out, valmode, seed);
}

/* LSD radix sort of 64-bit keys, 16-bit digits. */
static int radix_sort_u64(uint64_t *k, size_t n, int key_bits) {
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t) + 8);
    size_t *cnt = (size_t *)malloc(65536 * sizeof(size_t));
    if (!tmp || !cnt) { free(tmp); free(cnt); return -1; }
    uint64_t *a = k, *b = tmp;
    int passes = (key_bits + 15) / 16;
    for (int p = 0; p < passes; p++) {
        const int sh = 16 * p;
        memset(cnt, 0, 65536 * sizeof(size_t));
        for (size_t i = 0; i < n; i++) cnt[(a[i] >> sh) & 0xffff]++;
        size_t s = 0;
        for (int d = 0; d < 65536; d++) { size_t v = cnt[d]; cnt[d] = s; s += v; }
        for (size_t i = 0; i < n; i++) b[cnt[(a[i] >> sh) & 0xffff]++] = a[i];
        uint64_t *t = a; a = b; b = t;
    }
    if (a != k) memcpy(k, a, n * sizeof(uint64_t));
    free(tmp);
    free(cnt);
    return 0;
}

int tsg_gen_rmat(int scale, int edge_factor, double a, double b, double c, int symmetrize, int valmode,
                 unsigned long long seed, tsg_hcsr *out) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return TSG_ERR_ARG;
    const int n = 1 << scale;
    const size_t ne = (size_t)edge_factor << scale;
    const size_t nk = symmetrize ? 2 * ne : ne;
    uint64_t *keys = (uint64_t *)malloc(nk * sizeof(uint64_t));
    if (!keys) return TSG_ERR_NOMEM;
    const double ab = a + b, abc = a + b + c;
    for (size_t e = 0; e < ne; e++) {
        uint64_t u = 0, v = 0;
        for (int l = 0; l < scale; l++) {
            double r = u01(rng(seed, 5, (uint64_t)e * 64 + (uint64_t)l));
            int ub = 0, vb = 0;
            if (r < a) { }
            else if (r < ab) vb = 1;
            else if (r < abc) ub = 1;
            else { ub = 1; vb = 1; }
            u = (u << 1) | (uint64_t)ub;
            v = (v << 1) | (uint64_t)vb;
        }
        keys[e] = (u << 32) | v;
        if (symmetrize) keys[ne + e] = (v << 32) | u;
    }
    if (radix_sort_u64(keys, nk, 32 + scale)) { free(keys); return TSG_ERR_NOMEM; }
    size_t w = 0;
    for (size_t i = 0; i < nk; i++)
        if (w == 0 || keys[i] != keys[w - 1]) keys[w++] = keys[i];
    int st = alloc_csr(out, n, n, (long long)w);
    if (st) { free(keys); return st; }
    int row = 0;
    out->rowptr[0] = 0;
    for (size_t i = 0; i < w; i++) {
        int u = (int)(keys[i] >> 32);
        while (row < u) out->rowptr[++row] = (int)i;
        out->colidx[i] = (int)(keys[i] & 0xffffffffu);
        out->val[i] = 1.0;
    }
    while (row < n) out->rowptr[++row] = (int)w;
    free(keys);
    return tsg_hcsr_set_values(out, valmode == TSG_VAL_KEEP ? TSG_VAL_UNIFORM_01 : valmode, seed);
}

/* The batched workload (see tsg.h): draws of problem i come from stream
 * 100 of the batch seed, five per problem. */
int tsg_gen_batch_spec(int count, unsigned long long seed, int nmin, int nmax, tsg_batch_spec *out) {
    if (count < 0 || nmin < 1 || nmax < nmin || !out) return TSG_ERR_ARG;
    const double l0 = log((double)nmin), l1 = log((double)nmax);
    for (int i = 0; i < count; i++) {
        const uint64_t b = 5 * (uint64_t)i;
        tsg_batch_spec *s = &out[i];
        int n = (int)llround(exp(l0 + (l1 - l0) * u01(rng(seed, 100, b))));
        s->n = n < nmin ? nmin : (n > nmax ? nmax : n);
        s->family = u01(rng(seed, 100, b + 1)) < 0.5 ? 0 : 1;
        s->op = u01(rng(seed, 100, b + 2)) < 0.5 ? TSG_OP_A2 : TSG_OP_AAT;
        const double u = u01(rng(seed, 100, b + 3));
        s->mean_row = 4.0 + 28.0 * u;
        s->halfbw = 2 + (int)(39.0 * u);
        s->seed = rng(seed, 101, (uint64_t)i);
    }
    return TSG_OK;
}

int tsg_gen_batch_problem(const tsg_batch_spec *s, tsg_hcsr *out) {
    if (!s) return TSG_ERR_ARG;
    if (s->family == 0) return tsg_gen_er(s->n, s->mean_row, TSG_VAL_UNIFORM_PM1, s->seed, out);
    return tsg_gen_banded(s->n, s->halfbw, TSG_VAL_UNIFORM_PM1, s->seed, out);
}
