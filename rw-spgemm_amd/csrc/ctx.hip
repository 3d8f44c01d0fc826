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

// A block goes back to the arena it came from, else to the driver.
static void release_block(tsg_ctx *c, void *p, size_t sz) {
    if (tsg_arena_owns(c->arena, p)) tsg_arena_free(c->arena, p, sz);
    else hipFree(p);
}

extern "C" void tsg_ctx_destroy(tsg_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto &kv : c->free_blocks) release_block(c, kv.second, kv.first);
    for (auto &kv : c->live) release_block(c, kv.first, kv.second);
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
    hipStreamSynchronize(c->stream);  // no pending kernel still uses a cached block
    for (auto &kv : c->free_blocks) release_block(c, kv.second, kv.first);
    c->free_blocks.clear();
    c->bytes_cached = 0;
}

void tsg_ctx_set_arena(tsg_ctx *c, tsg_arena *a) { c->arena = a; }

// ------------------------------------------------------------------ arena
int tsg_arena_create(int device, size_t bytes, tsg_arena **out) {
    *out = nullptr;
    bytes = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
    TSG_HIP(hipSetDevice(device));
    tsg_arena *a = new tsg_arena();
    a->device = device;
    hipError_t e = hipMalloc((void **)&a->base, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        tsg_set_error("arena: device allocation of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
        delete a;
        return TSG_ERR_NOMEM;
    }
    a->size = bytes;
    a->free_by_off[0] = bytes;
    a->free_by_len.emplace(bytes, 0);
    a->bytes_free = bytes;
    *out = a;
    return TSG_OK;
}

void tsg_arena_destroy(tsg_arena *a) {
    if (!a) return;
    hipSetDevice(a->device);
    hipFree(a->base);
    delete a;
}

static void arena_erase_len(tsg_arena *a, size_t len, size_t off) {
    auto r = a->free_by_len.equal_range(len);
    for (auto it = r.first; it != r.second; ++it)
        if (it->second == off) {
            a->free_by_len.erase(it);
            return;
        }
}

void *tsg_arena_alloc(tsg_arena *a, size_t bytes) {
    std::lock_guard<std::mutex> g(a->mu);
    auto it = a->free_by_len.lower_bound(bytes);
    if (it == a->free_by_len.end()) return nullptr;
    const size_t len = it->first, off = it->second;
    a->free_by_len.erase(it);
    a->free_by_off.erase(off);
    if (len > bytes) {
        a->free_by_off[off + bytes] = len - bytes;
        a->free_by_len.emplace(len - bytes, off + bytes);
    }
    a->bytes_free -= bytes;
    return a->base + off;
}

void tsg_arena_free(tsg_arena *a, void *p, size_t bytes) {
    std::lock_guard<std::mutex> g(a->mu);
    size_t off = (size_t)((char *)p - a->base), len = bytes;
    a->bytes_free += bytes;
    auto nx = a->free_by_off.lower_bound(off);
    if (nx != a->free_by_off.end() && nx->first == off + len) {  // merge the next piece
        arena_erase_len(a, nx->second, nx->first);
        len += nx->second;
        nx = a->free_by_off.erase(nx);
    }
    if (nx != a->free_by_off.begin()) {  // merge the previous piece
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            arena_erase_len(a, pv->second, pv->first);
            off = pv->first;
            len += pv->second;
            a->free_by_off.erase(pv);
        }
    }
    a->free_by_off[off] = len;
    a->free_by_len.emplace(len, off);
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
    if (c->arena) {
        *p = tsg_arena_alloc(c->arena, sz);
        if (!*p && !c->free_blocks.empty()) {  // give the cache back and retry
            tsg_ctx_trim(c);
            *p = tsg_arena_alloc(c->arena, sz);
        }
        if (*p) {
            c->live[*p] = sz;
            c->bytes_live += sz;
            return TSG_OK;
        }
        // the arena is full: fall back to the driver
    }
    hipError_t e = hipMalloc(p, sz);
    if (e != hipSuccess) {
        // release the cache and retry once
        tsg_ctx_trim(c);
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

size_t tsg_alloc_size(const tsg_ctx *c, const void *p) {
    auto it = c->live.find(const_cast<void *>(p));
    return it == c->live.end() ? 0 : it->second;
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
