// scan.hip -- device exclusive prefix sums (replaces the thrust scan the
// reference calls in src/utils_cuda_scan.h:268-332).
//
// Reduce-then-scan in three launches, no inter-block waiting:
//   1. each 256-thread block sums a tile of 4096 counts (int64 partial)
//   2. one block scans the partials (int64), writes the grand total
//   3. each block rescans its tile and adds its partial offset
// Work is one read in pass 1 plus one read and one write in pass 3.
#include "tsg_internal.h"

#define SCAN_THREADS 256
#define SCAN_ITEMS 16
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_reduce(const int *__restrict__ in, long long n, long long *__restrict__ partial) {
    const long long base = (long long)blockIdx.x * SCAN_TILE;
    long long s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        long long i = base + (long long)j * SCAN_THREADS + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = tsg_wave_sum(s);
    __shared__ long long ws[SCAN_THREADS / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < SCAN_THREADS / 64; w++) t += ws[w];
        partial[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of nb int64 partials in place; total -> partial[nb].
__global__ void __launch_bounds__(1024)
k_scan_partials(long long *partial, long long nb) {
    __shared__ long long ws[16];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long long b0 = 0; b0 < nb; b0 += 1024) {
        long long i = b0 + threadIdx.x;
        long long v = i < nb ? partial[i] : 0;
        long long inc = tsg_wave_incl_scan(v);
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 63) ws[w] = inc;
        __syncthreads();
        long long woff = 0;
        for (int k = 0; k < w; k++) woff += ws[k];
        long long excl = carry + woff + inc - v;
        __syncthreads();
        if (i < nb) partial[i] = excl;
        if (threadIdx.x == 1023) carry = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

// Each thread rescans its 16 consecutive counts; a whole, 16-byte aligned
// run moves as four int4 loads and stores (the scalar form strides the
// lanes 64 bytes apart and splits every cache line over 16 instructions).
__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_down(const int *in, long long n, const long long *__restrict__ partial, int *out) {
    static_assert(SCAN_ITEMS == 16, "four int4 per thread");
    const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    const bool vec = base + SCAN_ITEMS <= n && !(((uintptr_t)in | (uintptr_t)out) & 15);
    int v[SCAN_ITEMS];
    if (vec) {
        const int4 *in4 = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int4 t = in4[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++) v[j] = base + j < n ? in[base + j] : 0;
    }
    long long s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) s += v[j];
    long long inc = tsg_wave_incl_scan(s);
    __shared__ long long ws[SCAN_THREADS / 64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    long long off = partial[blockIdx.x] + inc - s;
    for (int k = 0; k < w; k++) off += ws[k];
    int o[SCAN_ITEMS];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        o[j] = (int)off;
        off += v[j];
    }
    if (vec) {
        int4 *out4 = reinterpret_cast<int4 *>(out + base);
#pragma unroll
        for (int q = 0; q < 4; q++) out4[q] = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++)
            if (base + j < n) out[base + j] = o[j];
    }
}

__global__ void k_copy_total_i32(const long long *partial, long long nb, int *out, long long n) {
    long long t = partial[nb];
    out[n] = t > 0x7fffffffLL ? 0x7fffffff : (int)t;
}

long long tsg_scan_partial_count(long long n) {
    return (n <= 0 ? 1 : (n + SCAN_TILE - 1) / SCAN_TILE) + 1;
}

int tsg_scan_i32_enqueue(hipStream_t stream, const int *in, int *out, long long n, long long *partial) {
    if (n < 0) { tsg_set_error("tsg_scan_i32: n < 0"); return TSG_ERR_ARG; }
    const long long nb = tsg_scan_partial_count(n) - 1;
    // with n == 0 the single reduce block writes partial[0] = 0
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, stream, in, n, partial);
    TSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, stream, partial, nb);
    TSG_LAUNCH_CHECK();
    if (n > 0) {
        hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, stream, in, n,
                           (const long long *)partial, out);
        TSG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_copy_total_i32, dim3(1), dim3(1), 0, stream, (const long long *)partial, nb, out, n);
    TSG_LAUNCH_CHECK();
    return TSG_OK;
}

int tsg_scan_i32(tsg_ctx *ctx, const int *in, int *out, long long n, long long *total) {
    if (n < 0) { tsg_set_error("tsg_scan_i32: n < 0"); return TSG_ERR_ARG; }
    const long long np = tsg_scan_partial_count(n);
    void *scr;
    TSG_TRY(tsg_scratch(ctx, (size_t)np * sizeof(long long), &scr));
    long long *partial = (long long *)scr;
    TSG_TRY(tsg_scan_i32_enqueue(ctx->stream, in, out, n, partial));
    if (total) {
        TSG_TRY(tsg_read_scalar_i64(ctx, partial + np - 1, 1, total));
        if (*total > 0x7fffffffLL) {
            tsg_set_error("prefix sum total %lld exceeds int32", *total);
            return TSG_ERR_OVERFLOW;
        }
    }
    return TSG_OK;
}

// Plain streaming copy (16 B per lane per iteration): the reference point
// for reading hardware byte counters against a known byte count.
__global__ void __launch_bounds__(256)
k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long long n16) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n16; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

extern "C" int tsg_copy_kernel(tsg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) {
        tsg_set_error("tsg_copy_kernel: 16-byte aligned sizes and pointers only");
        return TSG_ERR_ARG;
    }
    const long long n16 = (long long)(bytes / 16);
    if (n16 > 0) {
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::min<long long>((n16 + 255) / 256, 8192)), dim3(256), 0,
                           ctx->stream, (const uint4 *)src, (uint4 *)dst, n16);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

__global__ void __launch_bounds__(256)
k_reduce_u64(const unsigned long long *__restrict__ in, long long n, unsigned long long *out) {
    unsigned long long s = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) s += in[i];
    s = tsg_wave_sum(s);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

int tsg_reduce_u64(tsg_ctx *ctx, const unsigned long long *in, long long n, unsigned long long *out) {
    void *scr;
    TSG_TRY(tsg_scratch(ctx, 64, &scr));
    TSG_HIP(hipMemsetAsync(scr, 0, 8, ctx->stream));
    if (n > 0) {
        unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_reduce_u64, dim3(nb), dim3(256), 0, ctx->stream, in, n, (unsigned long long *)scr);
        TSG_LAUNCH_CHECK();
    }
    long long t;
    TSG_TRY(tsg_read_scalar_i64(ctx, scr, 1, &t));
    *out = (unsigned long long)t;
    return TSG_OK;
}
