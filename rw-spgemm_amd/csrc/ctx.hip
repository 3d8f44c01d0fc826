// ctx.hip -- context, caching device allocator, error strings, host<->device
// CSR transfers.
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "tsg_internal.h"

static thread_local char g_err[1024] = "";

void tsg_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *tsg_last_error(void) { return g_err; }

extern "C" const char *tsg_build_info(void) {
    return "tilespgemm-mi355x: gfx950 HIP kernels, wave64, -O3 -ffp-contract=off, "
           "FP64 fma accumulation in ascending inner index";
}

extern "C" int tsg_ctx_create(int device, tsg_ctx **out) {
    if (!out) { tsg_set_error("tsg_ctx_create: out is NULL"); return TSG_ERR_ARG; }
    *out = nullptr;
    int ndev = 0;
    TSG_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        tsg_set_error("tsg_ctx_create: device %d out of range (%d devices)", device, ndev);
        return TSG_ERR_ARG;
    }
    TSG_HIP(hipSetDevice(device));
    tsg_ctx *c = new tsg_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        tsg_set_error("hipStreamCreate: %s", hipGetErrorString(e));
        delete c;
        return TSG_ERR_HIP;
    }
    e = hipHostMalloc((void **)&c->pinned, 64 * sizeof(long long), hipHostMallocDefault);
    if (e != hipSuccess) {
        tsg_set_error("hipHostMalloc: %s", hipGetErrorString(e));
        hipStreamDestroy(c->stream);
        delete c;
        return TSG_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->num_cu = prop.multiProcessorCount;
        c->lds_per_block = prop.sharedMemPerBlock;
    }
    *out = c;
    return TSG_OK;
}

extern "C" void tsg_ctx_destroy(tsg_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto &kv : c->free_blocks) hipFree(kv.second);
    for (auto &kv : c->live) hipFree(kv.first);
    if (c->scratch) hipFree(c->scratch);
    if (c->pinned) hipHostFree(c->pinned);
    hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int tsg_ctx_sync(tsg_ctx *c) {
    TSG_HIP(hipStreamSynchronize(c->stream));
    return TSG_OK;
}

extern "C" void *tsg_ctx_stream(tsg_ctx *c) { return (void *)c->stream; }
extern "C" long long tsg_ctx_bytes_in_use(tsg_ctx *c) { return c->bytes_live; }
extern "C" long long tsg_ctx_bytes_cached(tsg_ctx *c) { return c->bytes_cached; }

extern "C" void tsg_ctx_trim(tsg_ctx *c) {
    hipStreamSynchronize(c->stream);
    for (auto &kv : c->free_blocks) hipFree(kv.second);
    c->free_blocks.clear();
    c->bytes_cached = 0;
}

// Allocation sizes are rounded to 512 B (small) or 2 MiB (large) classes;
// a cached block is reused when it is at most 25% larger than asked.  All
// work of a context is ordered on its one stream, so a block freed by an
// earlier kernel can be handed to a later one without synchronisation.
static size_t round_size(size_t b) {
    if (b == 0) b = 1;
    if (b < (1u << 20)) return (b + 511) & ~size_t(511);
    const size_t g = size_t(2) << 20;
    return (b + g - 1) / g * g;
}

int tsg_alloc(tsg_ctx *c, void **p, size_t bytes) {
    *p = nullptr;
    size_t sz = round_size(bytes);
    auto it = c->free_blocks.lower_bound(sz);
    if (it != c->free_blocks.end() && it->first <= sz + sz / 4) {
        *p = it->second;
        c->live[*p] = it->first;
        c->bytes_live += it->first;
        c->bytes_cached -= it->first;
        c->free_blocks.erase(it);
        return TSG_OK;
    }
    hipError_t e = hipMalloc(p, sz);
    if (e != hipSuccess) {
        // release the cache and retry once
        hipStreamSynchronize(c->stream);
        for (auto &kv : c->free_blocks) hipFree(kv.second);
        c->free_blocks.clear();
        c->bytes_cached = 0;
        (void)hipGetLastError();
        e = hipMalloc(p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            tsg_set_error("device allocation of %zu bytes failed (%s); %lld bytes live",
                          sz, hipGetErrorString(e), c->bytes_live);
            return TSG_ERR_NOMEM;
        }
    }
    c->live[*p] = sz;
    c->bytes_live += sz;
    return TSG_OK;
}

void tsg_free(tsg_ctx *c, void *p) {
    if (!p) return;
    auto it = c->live.find(p);
    if (it == c->live.end()) return;
    c->free_blocks.emplace(it->second, p);
    c->bytes_live -= it->second;
    c->bytes_cached += it->second;
    c->live.erase(it);
}

int tsg_scratch(tsg_ctx *c, size_t bytes, void **p) {
    if (bytes > c->scratch_bytes) {
        if (c->scratch) {
            TSG_HIP(hipStreamSynchronize(c->stream));
            TSG_HIP(hipFree(c->scratch));
            c->scratch = nullptr;
        }
        size_t sz = round_size(bytes < (1u << 20) ? (1u << 20) : bytes);
        hipError_t e = hipMalloc(&c->scratch, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            c->scratch = nullptr;
            c->scratch_bytes = 0;
            tsg_set_error("scratch allocation of %zu bytes failed", sz);
            return TSG_ERR_NOMEM;
        }
        c->scratch_bytes = sz;
    }
    *p = c->scratch;
    return TSG_OK;
}

int tsg_read_scalar_i64(tsg_ctx *c, const void *dptr, int is64, long long *out) {
    TSG_HIP(hipMemcpyAsync(c->pinned, dptr, is64 ? 8 : 4, hipMemcpyDeviceToHost, c->stream));
    TSG_HIP(hipStreamSynchronize(c->stream));
    *out = is64 ? c->pinned[0] : (long long)((int *)c->pinned)[0];
    return TSG_OK;
}

// ------------------------------------------------------------------ host CSR
extern "C" void tsg_hcsr_free(tsg_hcsr *h) {
    if (!h) return;
    free(h->rowptr);
    free(h->colidx);
    free(h->val);
    h->rowptr = nullptr;
    h->colidx = nullptr;
    h->val = nullptr;
}

extern "C" int tsg_dcsr_upload(tsg_ctx *c, const tsg_hcsr *h, tsg_dcsr *d) {
    memset(d, 0, sizeof(*d));
    if (!h || h->m < 0 || h->n < 0 || h->nnz < 0 || !h->rowptr) {
        tsg_set_error("tsg_dcsr_upload: bad host CSR");
        return TSG_ERR_ARG;
    }
    d->m = h->m;
    d->n = h->n;
    d->nnz = h->nnz;
    TSG_TRY(tsg_alloc_t(c, &d->rowptr, (size_t)h->m + 1));
    TSG_TRY(tsg_alloc_t(c, &d->colidx, (size_t)h->nnz));
    TSG_TRY(tsg_alloc_t(c, &d->val, (size_t)h->nnz));
    TSG_HIP(hipMemcpyAsync(d->rowptr, h->rowptr, ((size_t)h->m + 1) * 4, hipMemcpyHostToDevice, c->stream));
    if (h->nnz) {
        TSG_HIP(hipMemcpyAsync(d->colidx, h->colidx, (size_t)h->nnz * 4, hipMemcpyHostToDevice, c->stream));
        TSG_HIP(hipMemcpyAsync(d->val, h->val, (size_t)h->nnz * 8, hipMemcpyHostToDevice, c->stream));
    }
    TSG_HIP(hipStreamSynchronize(c->stream));
    return TSG_OK;
}

extern "C" int tsg_dcsr_download(tsg_ctx *c, const tsg_dcsr *d, tsg_hcsr *h) {
    memset(h, 0, sizeof(*h));
    h->m = d->m;
    h->n = d->n;
    h->nnz = d->nnz;
    h->rowptr = (int *)malloc(((size_t)d->m + 1) * 4);
    h->colidx = (int *)malloc((size_t)(d->nnz ? d->nnz : 1) * 4);
    h->val = (double *)malloc((size_t)(d->nnz ? d->nnz : 1) * 8);
    if (!h->rowptr || !h->colidx || !h->val) {
        tsg_hcsr_free(h);
        tsg_set_error("tsg_dcsr_download: host allocation failed");
        return TSG_ERR_NOMEM;
    }
    TSG_HIP(hipMemcpyAsync(h->rowptr, d->rowptr, ((size_t)d->m + 1) * 4, hipMemcpyDeviceToHost, c->stream));
    if (d->nnz) {
        TSG_HIP(hipMemcpyAsync(h->colidx, d->colidx, (size_t)d->nnz * 4, hipMemcpyDeviceToHost, c->stream));
        TSG_HIP(hipMemcpyAsync(h->val, d->val, (size_t)d->nnz * 8, hipMemcpyDeviceToHost, c->stream));
    }
    TSG_HIP(hipStreamSynchronize(c->stream));
    return TSG_OK;
}

extern "C" void tsg_dcsr_free(tsg_ctx *c, tsg_dcsr *d) {
    if (!d) return;
    tsg_free(c, d->rowptr);
    tsg_free(c, d->colidx);
    tsg_free(c, d->val);
    memset(d, 0, sizeof(*d));
}
