// tsg_internal.h -- shared internals of the gfx950 TileSpGEMM library:
// context, caching device allocator, error plumbing and wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tsg.h"

#define TSG_TILE 16

// ------------------------------------------------------------------ errors
void tsg_set_error(const char *fmt, ...);

#define TSG_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            tsg_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,         \
                          hipGetErrorString(e_));                              \
            return TSG_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

#define TSG_TRY(expr)                                                          \
    do {                                                                       \
        int s_ = (expr);                                                       \
        if (s_ != TSG_OK) return s_;                                           \
    } while (0)

#define TSG_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            tsg_set_error("%s:%d kernel launch -> %s", __FILE__, __LINE__,     \
                          hipGetErrorString(e_));                              \
            return TSG_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

// ------------------------------------------------------------------ arena
// One device allocation handed out in pieces (best fit; released pieces
// coalesce with free neighbours); shared by contexts on several threads.
// Memory returned to an arena stays mapped, so a workload that keeps
// allocating and releasing large blocks does not go back to the driver.
struct tsg_arena {
    int device = 0;
    char *base = nullptr;
    size_t size = 0;
    std::mutex mu;
    std::map<size_t, size_t> free_by_off;       // offset -> length
    std::multimap<size_t, size_t> free_by_len;  // length -> offset
    size_t bytes_free = 0;
};
int tsg_arena_create(int device, size_t bytes, tsg_arena **out);
void tsg_arena_destroy(tsg_arena *a);
void *tsg_arena_alloc(tsg_arena *a, size_t bytes);  // nullptr when no free piece fits
void tsg_arena_free(tsg_arena *a, void *p, size_t bytes);
inline bool tsg_arena_owns(const tsg_arena *a, const void *p) {
    return a && (const char *)p >= a->base && (const char *)p < a->base + a->size;
}

// ------------------------------------------------------------------ context
struct tsg_ctx {
    int device;
    hipStream_t stream;
    // caching allocator: size -> blocks (stream-ordered reuse on `stream`)
    std::multimap<size_t, void *> free_blocks;
    std::map<void *, size_t> live;
    long long bytes_live = 0;
    long long bytes_cached = 0;
    // small pinned host buffer for scalar read-backs
    long long *pinned = nullptr;
    // persistent small device scratch (scan partials, counters)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int num_cu = 0;
    size_t lds_per_block = 0;
    // when set, blocks come from this arena (hipMalloc only if it is full)
    tsg_arena *arena = nullptr;
};
// Routes ctx's allocations through an arena; call before it allocates.
void tsg_ctx_set_arena(tsg_ctx *ctx, tsg_arena *a);

int tsg_alloc(tsg_ctx *ctx, void **p, size_t bytes);
void tsg_free(tsg_ctx *ctx, void *p);
template <typename T>
static inline int tsg_alloc_t(tsg_ctx *ctx, T **p, size_t count) {
    return tsg_alloc(ctx, (void **)p, count * sizeof(T));
}
int tsg_scratch(tsg_ctx *ctx, size_t bytes, void **p);
int tsg_read_scalar_i64(tsg_ctx *ctx, const void *dptr, int is64, long long *out);
// Bytes the allocator reserved for a live block of ctx (0 if none).
size_t tsg_alloc_size(const tsg_ctx *ctx, const void *p);

// ------------------------------------------------------------------ scans
// Exclusive scan of n non-negative int32 counts into out[0..n] (out[n] =
// total).  `in` may alias `out`.  total (may be NULL) is read back to the
// host; TSG_ERR_OVERFLOW when the total exceeds INT32_MAX.
int tsg_scan_i32(tsg_ctx *ctx, const int *in, int *out, long long n, long long *total);
// Enqueue-only form (no host read-back, no memset: graph-capturable).
// `partial` holds tsg_scan_partial_count(n) int64 values; the int64 total
// is left in partial[tsg_scan_partial_count(n) - 1].
long long tsg_scan_partial_count(long long n);
int tsg_scan_i32_enqueue(hipStream_t stream, const int *in, int *out, long long n, long long *partial);
// Same for unsigned-char counts of 16-row tiles is done inside the kernels.
// Sum of n int64 values (device) -> host.
int tsg_reduce_u64(tsg_ctx *ctx, const unsigned long long *in, long long n,
                   unsigned long long *out);

// ------------------------------------------------------------------ spgemm
// Decisions of one SpGEMM that depend only on the operand patterns (sizes
// of C, which step-2 / step-3 kernels run, the LDS layout of the row-wise
// step 3).  A plan made by tsg_spgemm_planned lets tsg_spgemm_enqueue
// recompute the same product into the same C buffers with no host
// read-back, allocation or memset (so the sequence can be captured in a
// HIP graph).
struct tsg_spgemm_plan {
    long long numtileC, nnzC;
    int maxrow;     // C tiles of the longest tile-row
    int all;        // every tile-row through the tile-parallel kernels
    int s3_rows;    // some tile-row runs the row-wise step 3
    int s3_unfit;   // some tile-row runs the tile-parallel step 3
    int dense_min;  // dense accumulator threshold of the tile-parallel step 3
    int s3_elem;    // the tile-rows the row-wise kernel cannot hold go element-wise
    int elem_all;   // every tile-row goes element-wise
    int g3[6];      // row-wise step-3 LDS layout maxima
    int s1w;        // step-1 bitmap window, 32-bit words per wave
    int s2u;        // most C tiles of a step-2 tile-pair segment (LDS sizing)
    int s2map;      // bytes of the step-2 column -> slot map (0: binary search)
    int s3pk;       // element step 3 packs hash keys and positions (narrow tile-rows)
};
// symbolic = true runs steps 1-2 and the step-3 plan only (C's values and
// in-tile columns stay unallocated); stats (when given) then holds C's exact
// tile and nonzero counts and the numeric step's heavy-row scratch bytes.
// On any error every temporary and every array of C allocated so far is
// freed (C is left zeroed).
int tsg_spgemm_planned(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, const tsg_tune *tune,
                       tsg_dtiles *C, tsg_stats *stats, tsg_spgemm_plan *plan_out, bool symbolic = false);
// tsg_spgemm_tile_rows with the symbolic option above.
int tsg_spgemm_tile_rows_ex(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, int r0, int r1,
                            const tsg_tune *tune, tsg_dtiles *C, tsg_stats *stats, bool symbolic);
// Intermediate products of each A tile-row (device u64[A->tilem]).
int tsg_tile_row_products(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, unsigned long long *d_out);
// Device scratch bytes tsg_spgemm_enqueue needs for this C and B (B may be
// NULL: no room for B's tile-row bitmaps, counted for a product with B).
size_t tsg_spgemm_work_bytes(const tsg_dtiles *C, const tsg_dtiles *B, int s1w);
// Step-1 bitmap window (32-bit words per wave) for products with this B.
int tsg_s1_words(const tsg_dtiles *B, const tsg_tune &t);
int tsg_spgemm_enqueue(hipStream_t s, const tsg_dtiles *A, const tsg_dtiles *B, tsg_dtiles *C,
                       const tsg_spgemm_plan &P, void *work);

// ------------------------------------------------------------------ conversions
int tsg_tile2csr_view(tsg_ctx *ctx, const tsg_dtiles *t, int jbase, tsg_dcsr *out);
// Block-diagonal packing of tile matrices on 16-row/column boundaries.
int tsg_tiles_pack(tsg_ctx *ctx, const tsg_dtiles *const *parts, int count, tsg_dtiles *out);

// ------------------------------------------------------------------ sort
// Stable ascending sort of (key,value) pairs, 32-bit keys, using the
// library's own LSD radix sort.  Keys must be < (1u << key_bits).
int tsg_sort_pairs_u32(tsg_ctx *ctx, unsigned *keys, int *vals, long long n,
                       int key_bits);
int tsg_sort_pairs_u64(tsg_ctx *ctx, unsigned long long *keys, int *vals, long long n,
                       int key_bits);

// ------------------------------------------------------------------ device helpers
__device__ __forceinline__ void tsg_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int tsg_lane() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T tsg_wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T tsg_wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, 64); v = w < v ? w : v; }
    return v;
}
template <typename T>
__device__ __forceinline__ T tsg_wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, 64); v = w > v ? w : v; }
    return v;
}
// inclusive scan across the 64 lanes
template <typename T>
__device__ __forceinline__ T tsg_wave_incl_scan(T v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    return v;
}
// inclusive scan inside groups of 16 lanes
template <typename T>
__device__ __forceinline__ T tsg_group16_incl_scan(T v) {
    const int r = __lane_id() & 15;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        T w = __shfl_up(v, o, 16);
        if (r >= o) v += w;
    }
    return v;
}
// Value of lane q of this lane's group of 16 lanes (q a constant after
// unrolling): a DPP row broadcast, a VALU move with no LDS traffic, where
// __shfl would issue a ds_bpermute.  Every lane of the group must be active.
#define TSG_RB_CASE(Q) case Q: return __builtin_amdgcn_mov_dpp(v, 0x150 + Q, 0xF, 0xF, false);
__device__ __forceinline__ int tsg_row16_bcast(int v, int q) {
    switch (q) {
        TSG_RB_CASE(0) TSG_RB_CASE(1) TSG_RB_CASE(2) TSG_RB_CASE(3)
        TSG_RB_CASE(4) TSG_RB_CASE(5) TSG_RB_CASE(6) TSG_RB_CASE(7)
        TSG_RB_CASE(8) TSG_RB_CASE(9) TSG_RB_CASE(10) TSG_RB_CASE(11)
        TSG_RB_CASE(12) TSG_RB_CASE(13) TSG_RB_CASE(14) default: break;
    }
    return __builtin_amdgcn_mov_dpp(v, 0x150 + 15, 0xF, 0xF, false);
}
#undef TSG_RB_CASE
__device__ __forceinline__ double tsg_row16_bcast(double v, int q) {
    const long long b = __double_as_longlong(v);
    const int lo = tsg_row16_bcast((int)(unsigned)b, q), hi = tsg_row16_bcast((int)(b >> 32), q);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <typename T>
__device__ __forceinline__ T tsg_group16_sum(T v) {
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
    return v;
}

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X; remap so
// that each XCD (and its L2) works on one contiguous range of block ids.
__device__ __forceinline__ unsigned tsg_xcd_block(unsigned bid, unsigned nblocks) {
    const unsigned nx = 8;
    const unsigned per = nblocks / nx;      // full rounds
    const unsigned rem = nblocks % nx;
    const unsigned xcd = bid % nx, k = bid / nx;
    // XCD x owns per (+1 if x < rem) consecutive logical ids
    const unsigned base = xcd * per + (xcd < rem ? xcd : rem);
    return (k < per + (xcd < rem ? 1u : 0u)) ? base + k : bid;
}

static inline unsigned tsg_div_up(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
int tsg_validate_csr(tsg_ctx *ctx, const tsg_dcsr *a);
// sorted = 0: row pointers and column range only (repeated or unsorted
// columns allowed, as mmio_allinone leaves them).
int tsg_validate_csr_mode(tsg_ctx *ctx, const tsg_dcsr *a, int sorted);
