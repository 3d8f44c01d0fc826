// sort.hip -- stable LSD radix sort of (key, int value) pairs.
//
// Used where the reference calls thrust::sort / sort_by_key
// (src/reorder.h:652,721,770,2495) and for the stable CSR transpose.
// One pass per 8-bit digit:
//   hist    : each block counts the digits of its 2048-key tile (LDS)
//   scan    : digit-major exclusive scan of the block histograms
//   scatter : each block ranks its keys stably -- tile order is
//             (iteration, thread) so keys are ranked iteration by
//             iteration; inside a wave64 the lanes that share a digit are
//             found with 8 ballots, across waves with LDS counters.
#include "tsg_internal.h"

#define RS_THREADS 256
#define RS_ITEMS 8
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_BITS 8
#define RS_BUCKETS 256

template <typename K>
__global__ void __launch_bounds__(RS_THREADS)
k_rs_hist(const K *__restrict__ keys, long long n, int shift, int nblocks, int *__restrict__ hist) {
    __shared__ int h[RS_BUCKETS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        long long i = base + (long long)it * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & (RS_BUCKETS - 1)], 1);
    }
    __syncthreads();
    hist[(long long)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <typename K>
__global__ void __launch_bounds__(RS_THREADS)
k_rs_scatter(const K *__restrict__ kin, const int *__restrict__ vin, long long n, int shift, int nblocks,
             const int *__restrict__ offs, K *__restrict__ kout, int *__restrict__ vout) {
    __shared__ int running[RS_BUCKETS];
    __shared__ int wcnt[RS_THREADS / 64][RS_BUCKETS];
    __shared__ int base_off[RS_BUCKETS];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    running[tid] = 0;
    base_off[tid] = offs[(long long)tid * nblocks + blockIdx.x];
    for (int k = 0; k < RS_THREADS / 64; k++) wcnt[k][tid] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * RS_TILE;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int it = 0; it < RS_ITEMS; it++) {
        long long i = base + (long long)it * RS_THREADS + tid;
        const bool valid = i < n;
        K key = valid ? kin[i] : (K)0;
        int val = valid ? vin[i] : 0;
        unsigned d = (unsigned)(key >> shift) & (RS_BUCKETS - 1);
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RS_BITS; b++) {
            unsigned long long bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const int wrank = __popcll(peers & lt);
        const bool leader = valid && wrank == 0;
        if (leader) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int off = base_off[d] + running[d] + wrank;
            for (int k = 0; k < w; k++) off += wcnt[k][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        int add = 0;
        for (int k = 0; k < RS_THREADS / 64; k++) { add += wcnt[k][tid]; wcnt[k][tid] = 0; }
        running[tid] += add;
        __syncthreads();
    }
}

template <typename K>
static int sort_pairs(tsg_ctx *ctx, K *keys, int *vals, long long n, int key_bits) {
    if (n <= 1) return TSG_OK;
    if (n > 0x7fffffffLL) { tsg_set_error("radix sort: n too large"); return TSG_ERR_OVERFLOW; }
    const int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
    const int passes = (key_bits + RS_BITS - 1) / RS_BITS;
    K *k2 = nullptr;
    int *v2 = nullptr, *hist = nullptr;
    TSG_TRY(tsg_alloc_t(ctx, &k2, (size_t)n));
    TSG_TRY(tsg_alloc_t(ctx, &v2, (size_t)n));
    TSG_TRY(tsg_alloc_t(ctx, &hist, (size_t)RS_BUCKETS * nblocks + 1));
    K *ka = keys, *kb = k2;
    int *va = vals, *vb = v2;
    for (int p = 0; p < passes; p++) {
        const int shift = p * RS_BITS;
        hipLaunchKernelGGL(k_rs_hist<K>, dim3(nblocks), dim3(RS_THREADS), 0, ctx->stream, ka, n, shift, nblocks, hist);
        TSG_LAUNCH_CHECK();
        TSG_TRY(tsg_scan_i32(ctx, hist, hist, (long long)RS_BUCKETS * nblocks, nullptr));
        hipLaunchKernelGGL(k_rs_scatter<K>, dim3(nblocks), dim3(RS_THREADS), 0, ctx->stream, ka, va, n, shift, nblocks,
                           hist, kb, vb);
        TSG_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        TSG_HIP(hipMemcpyAsync(keys, ka, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream));
        TSG_HIP(hipMemcpyAsync(vals, va, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
    }
    tsg_free(ctx, k2);
    tsg_free(ctx, v2);
    tsg_free(ctx, hist);
    return TSG_OK;
}

int tsg_sort_pairs_u32(tsg_ctx *ctx, unsigned *keys, int *vals, long long n, int key_bits) {
    return sort_pairs<unsigned>(ctx, keys, vals, n, key_bits < 1 ? 1 : (key_bits > 32 ? 32 : key_bits));
}
int tsg_sort_pairs_u64(tsg_ctx *ctx, unsigned long long *keys, int *vals, long long n, int key_bits) {
    return sort_pairs<unsigned long long>(ctx, keys, vals, n, key_bits < 1 ? 1 : (key_bits > 64 ? 64 : key_bits));
}
