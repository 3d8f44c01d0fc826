// spgemm.hip -- the TileSpGEMM pipeline on gfx950 (wave64, LDS, FP64 FMA).
//
// Step 1 (tile-level symbolic, src/tilespgemm-cuda.h:257-369 and the nsparse
//   hash path src/spgemm_nsparse_kernel.h:138-1511): for each A tile-row I the
//   set of C tile columns  U_{K in A(I,:)} B(K,:)  -- one wave per tile-row,
//   the (K, J) candidates flattened across the 64 lanes and OR-ed into a
//   windowed LDS bitmap; count pass, device scan, then a fill pass emitting
//   the sorted tile columns.  Covers every tilen (no SPA/hash split).
// Step 2 (16x16 bitmask symbolic, src/tilespgemm-cuda.h:371-755):
//   k_step2_rows, one block per C tile-row, visits every (A tile, B tile)
//   pair of the row once (no intersection search) and ORs the pair's
//   boolean product rows into the C tile masks held in LDS; popcounts give
//   the u8 row offsets and tile nnz.  Tile-rows with more C tiles than LDS
//   holds use k_step2_tiles (one C tile per 16-lane group, pairs from a
//   sorted-list intersection, 64-bit ballots to keep them in K order).
// Step 3 (numeric, src/tilespgemm-cuda.h:1505-2608): k_step3_rows, one
//   block per C tile-row, accumulates the row's C values in LDS in their
//   final order; each C row belongs to one wave, which takes its tasks in
//   ascending K.  A tile K with several entries has A tile K and B tile-row
//   K staged into LDS (software pipelined) and the (B tile, A row) tasks
//   that meet nonempty B rows run compacted per wave; a single-entry A tile
//   is handled by its row's wave from registers.  Hypersparse and hub rows
//   use k_step3_elem / k_step3_heavy (B's row index); larger tile-rows
//   otherwise use k_step3 (one C tile per
//   16-lane group, lane r owns tile row r, dense or compact LDS row
//   accumulator).  No atomics anywhere: every C entry receives its
//   products in ascending inner index k, the order of the serial SPA
//   reference, so results are deterministic and bit-identical to it.  (The
//   reference uses FP64 atomicAdd, src/tilespgemm-cuda.h:1670.)
#include <string.h>

#include <mutex>

#include "tsg_internal.h"

#define S1_WAVES 4
#define S23_WAVES 4
#define TILES_PER_WAVE 4

extern "C" void tsg_tune_default(tsg_tune *t) {
    memset(t, 0, sizeof(*t));
    t->dense_acc_min_nnz = 64;
    t->s1_window_words = 0;
}

// ------------------------------------------------------------------ B tile-row bitmaps
// For step 1: the tile columns of B tile-row K as a bitmap of nw 32-bit
// words from absolute word F (bit b of word w <-> tile column 32 (F + w) + b),
// kept at bits + off when it is at most about twice the size of the row's
// column list (nw <= 2 * tiles + 2); desc[K] = {off, F, nw, 0} with nw = 0
// for a row that is empty or not kept.  off = 2 * Bptr[K] + 2 * K bounds
// every row's room without a scan.  Two launches (graph-capturable, every
// word of a kept row written once, no memset): k_brow_desc, one lane per
// tile-row, then k_brow_fill, one lane per B tile, so a hub tile-row's
// words are spread over as many waves as it has tiles.
__global__ void __launch_bounds__(256)
k_brow_desc(const int *__restrict__ Bptr, const int *__restrict__ Bcol, int tilem, int4 *__restrict__ desc,
            int *__restrict__ any) {
    const int K = blockIdx.x * 256 + threadIdx.x;
    int nw = 0;
    if (K < tilem) {
        const int b0 = Bptr[K], n = Bptr[K + 1] - b0;
        int F = 0;
        if (n > 0) {
            F = Bcol[b0] >> 5;
            nw = (Bcol[b0 + n - 1] >> 5) - F + 1;
            if (nw > 2 * n + 2) nw = 0;
        }
        desc[K] = make_int4(2 * b0 + 2 * K, F, nw, 0);
    }
    // any (may be NULL, zeroed beforehand): set when some row is kept; the
    // same value from every writer, and no store once it reads 1
    if (any && __ballot(nw > 0) && (threadIdx.x & 63) == 0 && *(volatile int *)any == 0) *(volatile int *)any = 1;
}

// A run is the tiles of one kept row in one bitmap word (consecutive in the
// row-major tile order).  Its last tile's lane writes the word (the run's
// bits OR-ed across lanes, plus the run's tiles before the wave's first,
// at most 31) and zeroes the empty words up to the row's next run.
__global__ void __launch_bounds__(256)
k_brow_fill(const int *__restrict__ Brow, const int *__restrict__ Bcol, long long numtile,
            const int4 *__restrict__ desc, unsigned *__restrict__ bits, const int *__restrict__ any) {
    const int lane = threadIdx.x & 63;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long t0 = t - lane;  // the wave's first tile
    if (t0 >= numtile) return;      // whole waves
    if (any && *any == 0) return;   // no row kept (k_brow_desc's flag, when given)
    const bool valid = t < numtile;
    int K = -1, w = -1, col = 0;
    int4 d = make_int4(0, 0, 0, 0);
    if (valid) {
        K = Brow[t];
        d = desc[K];
        col = Bcol[t];
        if (d.z > 0) w = (col >> 5) - d.y;
    }
    const bool kept = valid && d.z > 0;
    // the next tile's row and word (lane 63: loaded)
    int Kn = __shfl_down(K, 1, 64), wn = __shfl_down(w, 1, 64), cn = __shfl_down(col, 1, 64);
    if (lane == 63) {
        Kn = -1;
        if (t + 1 < numtile) {
            Kn = Brow[t + 1];
            cn = Bcol[t + 1];
            wn = Kn == K ? (cn >> 5) - d.y : -1;
        }
    }
    const bool last = kept && !(Kn == K && wn == w);
    // inclusive OR over the run's lanes up to this one
    unsigned v = kept ? 1u << (col & 31) : 0u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned vv = __shfl_up(v, o, 64);
        const int kk = __shfl_up(K, o, 64), ww = __shfl_up(w, o, 64);
        if (lane >= o && kk == K && ww == w) v |= vv;
    }
    // lane 0's run, read with every lane active (a shuffle inside the branch
    // below would not see lane 0 unless it is a run's last lane too)
    const int K0 = __shfl(K, 0, 64), w0 = __shfl(w, 0, 64);
    if (last) {
        if (K0 == K && w0 == w)  // the run began before this wave
            for (long long u = t0 - 1; u >= 0 && Brow[u] == K; u--) {
                const int c = Bcol[u];
                if ((c >> 5) - d.y != w) break;
                v |= 1u << (c & 31);
            }
        unsigned *row = bits + d.x;
        row[w] = v;
        const int wend = Kn == K ? wn : d.z;  // the row's next run, or its end
        for (int g = w + 1; g < wend; g++) row[g] = 0u;
    }
}

// ------------------------------------------------------------------ step 1
// Bdesc / Bbits (may be NULL): B tile-rows kept as bitmaps (k_brow_desc,
// k_brow_fill).  Their candidates are not enumerated: each lane ORs the
// bitmap words over its own window words, with no search and no LDS
// atomics.  Bany (may be NULL) is k_brow_desc's flag: when no row was kept,
// the descriptors are not read at all.
template <bool FILL, int BMW>
__global__ void __launch_bounds__(64 * S1_WAVES)
k_step1(const int *__restrict__ Aptr, const int *__restrict__ Acol, int tilem,
        const int *__restrict__ Bptr, const int *__restrict__ Bcol, const int4 *__restrict__ Bdesc,
        const unsigned *__restrict__ Bbits, const int *__restrict__ Bany, int *__restrict__ Cptr_or_cnt,
        int *__restrict__ Ccol, int *__restrict__ Crow, unsigned *__restrict__ bmstore, int *__restrict__ rowspan) {
    __shared__ unsigned s_bm[S1_WAVES][BMW];
    __shared__ int s_pre[S1_WAVES][64];
    __shared__ int s_b0[S1_WAVES][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = blockIdx.x * S1_WAVES + wave;
    if (I >= tilem) return;
    unsigned *bm = s_bm[wave];
    int *pre = s_pre[wave], *pb0 = s_b0[wave];
    const int a0 = Aptr[I], a1 = Aptr[I + 1];
    if (Bany && *Bany == 0) Bdesc = nullptr;  // wave-uniform

    // range of candidate tile columns: B rows are sorted
    int jmin = 0x7fffffff, jmax = -1;
    for (int p = a0 + lane; p < a1; p += 64) {
        const int K = Acol[p];
        const int b0 = Bptr[K], b1 = Bptr[K + 1];
        if (b1 > b0) {
            jmin = min(jmin, Bcol[b0]);
            jmax = max(jmax, Bcol[b1 - 1]);
        }
    }
    jmin = tsg_wave_min(jmin);
    jmax = tsg_wave_max(jmax);
    // windows start on a word boundary of absolute tile columns (the words
    // of B's tile-row bitmaps)
    const int jb = jmin & ~31;

    // a row spanning 2..64 windows (a periodic grid's boundary rows) first
    // finds the windows its candidates hit and visits only those
    constexpr int WSH = __builtin_ctz(32 * BMW);  // log2 of the columns per window
    unsigned long long wmask = ~0ull;
    if (jmax >= jmin && ((jmax - jb) >> WSH) >= 1 && ((jmax - jb) >> WSH) < 64) {
        unsigned wlo = 0u, whi = 0u;
        for (int c0 = a0; c0 < a1; c0 += 64) {
            const int p = c0 + lane;
            int b0 = 0, len = 0;
            if (p < a1) {
                const int K = Acol[p];
                b0 = Bptr[K];
                len = Bptr[K + 1] - b0;
                if (Bdesc) {
                    const int4 d = Bdesc[K];
                    if (d.z > 0) {  // a kept row marks the windows its words span
                        len = 0;
                        const int i0 = (max(d.y << 5, jb) - jb) >> WSH;
                        const int i1 = (min(((d.y + d.z) << 5) - 1, jmax) - jb) >> WSH;
                        const unsigned long long m = (i1 >= 63 ? ~0ull : (2ull << i1) - 1) & ~((1ull << i0) - 1);
                        wlo |= (unsigned)m;
                        whi |= (unsigned)(m >> 32);
                    }
                }
            }
            const int incl = tsg_wave_incl_scan(len);
            const int tot = __shfl(incl, 63, 64);
            pre[lane] = incl - len;
            pb0[lane] = b0;
            tsg_wave_sync();
            for (int c = lane; c < tot; c += 64) {
                int lo = 0;
#pragma unroll
                for (int st = 32; st > 0; st >>= 1)
                    if (pre[lo + st] <= c) lo += st;
                const int wi = (Bcol[pb0[lo] + (c - pre[lo])] - jb) >> WSH;
                if (wi < 32) wlo |= 1u << wi; else whi |= 1u << (wi - 32);
            }
            tsg_wave_sync();
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            wlo |= __shfl_xor(wlo, o, 64);
            whi |= __shfl_xor(whi, o, 64);
        }
        wmask = ((unsigned long long)whi << 32) | wlo;
    }

    int total = 0;
    const int base = FILL ? Cptr_or_cnt[I] : 0;
    // a row whose candidates fit one window keeps the count pass's bitmap
    // (bmstore, BMW words per tile-row) so the fill pass reads it back
    const bool kept = bmstore != nullptr && jmax >= jmin && jmax - jb < 32 * BMW;
    unsigned *keep = kept ? bmstore + (long long)I * BMW : nullptr;
    constexpr int V = BMW / 64;            // window words per lane: lane + 64 v
    constexpr int U = BMW == 256 ? 2 : 1;  // kept B tile-rows whose words are loaded together
    for (int w0 = jb; w0 <= jmax; w0 += 32 * BMW) {
        const int wi = (w0 - jb) >> WSH;
        if (wi < 64 && !((wmask >> wi) & 1ull)) continue;  // no candidate in this window
        const int nw = min(BMW, ((jmax - w0) >> 5) + 1);
        if (FILL && kept) {
            for (int k = lane; k < nw; k += 64) bm[k] = keep[k];
            tsg_wave_sync();
        } else {
            for (int k = lane; k < nw; k += 64) bm[k] = 0u;
            tsg_wave_sync();
            const int W0 = w0 >> 5;  // absolute word of the window's first word
            for (int c0 = a0; c0 < a1; c0 += 64) {
                const int p = c0 + lane;
                int b0 = 0, len = 0;
                int4 d = make_int4(0, 0, 0, 0);
                if (p < a1) {
                    const int K = Acol[p];
                    b0 = Bptr[K];
                    len = Bptr[K + 1] - b0;
                    if (Bdesc) {
                        d = Bdesc[K];
                        if (d.z > 0) len = 0;
                    }
                }
                // kept B tile-rows: window words [lo, hi) of each, U rows' loads
                // in flight; a lane ORs only its own words (lane + 64 v), so
                // no atomics, and LDS order within the wave keeps these
                // before the candidate atomics below
                unsigned long long dm = __ballot(d.z > 0);
                while (dm) {
                    unsigned t[U][V];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        int lo = 0, hi = 0, base = 0;
                        if (dm) {
                            const int j = __ffsll((long long)dm) - 1;
                            dm &= dm - 1;
                            const int dF = __builtin_amdgcn_readlane(d.y, j);
                            const int dn = __builtin_amdgcn_readlane(d.z, j);
                            lo = max(W0, dF) - W0;
                            hi = min(W0 + nw, dF + dn) - W0;
                            base = __builtin_amdgcn_readlane(d.x, j) + W0 - dF;
                        }
#pragma unroll
                        for (int v = 0; v < V; v++) {
                            const int k = 64 * v + lane;
                            t[u][v] = 0u;
                            if (64 * v + 64 > lo && 64 * v < hi && k >= lo && k < hi) t[u][v] = Bbits[base + k];
                        }
                    }
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        unsigned x = 0u;
#pragma unroll
                        for (int u = 0; u < U; u++) x |= t[u][v];
                        if (x) bm[64 * v + lane] |= x;
                    }
                }
                const int incl = tsg_wave_incl_scan(len);
                const int tot = __shfl(incl, 63, 64);
                pre[lane] = incl - len;
                pb0[lane] = b0;
                tsg_wave_sync();
                // 4 B tile columns per lane are loaded before their bits are set
                if (tot >= 64 * 32) {
                    // long B tile-rows (32+ tiles on average): the 64 lanes sweep
                    // one B tile-row at a time, no candidate search
                    const int na = min(64, a1 - c0);
                    for (int j = 0; j < na; j++) {
                        const int bj = pb0[j], lj = (j + 1 < na ? pre[j + 1] : tot) - pre[j];
                        for (int c = lane; c < lj; c += 256) {
                            int d[4];
#pragma unroll
                            for (int v = 0; v < 4; v++) d[v] = c + 64 * v < lj ? Bcol[bj + c + 64 * v] - w0 : -1;
#pragma unroll
                            for (int v = 0; v < 4; v++)
                                if (d[v] >= 0 && d[v] < 32 * BMW) atomicOr(&bm[d[v] >> 5], 1u << (d[v] & 31));
                        }
                    }
                } else {
                    for (int c = lane; c < tot; c += 256) {
                        int d[4];
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int cc = c + 64 * v;
                            d[v] = -1;
                            if (cc < tot) {
                                int lo = 0;
#pragma unroll
                                for (int st = 32; st > 0; st >>= 1)
                                    if (pre[lo + st] <= cc) lo += st;
                                d[v] = Bcol[pb0[lo] + (cc - pre[lo])] - w0;
                            }
                        }
#pragma unroll
                        for (int v = 0; v < 4; v++)
                            if (d[v] >= 0 && d[v] < 32 * BMW) atomicOr(&bm[d[v] >> 5], 1u << (d[v] & 31));
                    }
                }
                tsg_wave_sync();
            }
            if (!FILL && kept)
                for (int k = lane; k < nw; k += 64) keep[k] = bm[k];
        }
        if (!FILL) {
            int cnt = 0;
            for (int k = lane; k < nw; k += 64) cnt += __popc(bm[k]);
            total += tsg_wave_sum(cnt);
        } else {
            const int chunk = (nw + 63) >> 6;
            const int k0 = lane * chunk, k1 = min(nw, k0 + chunk);
            int cnt = 0;
            for (int k = k0; k < k1; k++) cnt += __popc(bm[k]);
            int pos = base + total + tsg_wave_incl_scan(cnt) - cnt;
            for (int k = k0; k < k1; k++) {
                unsigned x = bm[k];
                while (x) {
                    const int b = __ffs(x) - 1;
                    x &= x - 1;
                    Ccol[pos] = w0 + k * 32 + b;
                    Crow[pos] = I;
                    pos++;
                }
            }
            total += tsg_wave_sum(cnt);
        }
        tsg_wave_sync();
    }
    if (!FILL && lane == 0) {
        Cptr_or_cnt[I] = total;
        // the candidates are exactly C's tiles, so this is the row's C tile-column span
        if (rowspan) rowspan[I] = total > 0 ? jmax - jmin + 1 : 0;
    }
}

// ------------------------------------------------------------------ pairs
// Operands of steps 2 and 3.  B's tile data is in row-major tile order;
// its CSC index (Bcscptr/Bcscrow/Bcscid) lists each tile column's tiles by
// ascending tile row.
struct TileOperands {
    const int *Aptr, *Acol, *Annz;
    const unsigned char *Aptr16;
    const unsigned short *Amask;
    const double *Aval;
    const int *Bptr, *Bcol;
    const int *Bcscptr, *Bcscrow, *Bcscid, *Bnnz;
    const unsigned char *Bptr16;
    const unsigned short *Bmask;
    const double *Bval;
    // A's element row index (may be NULL): rows' entries in ascending k
    const int *Arowp, *Acolx;
    const double *Avalx;
};

// Index p in [0, len) with list[p] == key, or -1 (list ascending).
__device__ __forceinline__ int find_sorted(const int *__restrict__ list, int len, int key) {
    int lo = 0, hi = len - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = list[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

#define PAIR_LDS_MAX 64  // both tile lists staged in LDS when each is this short

// Calls body(ta, tb) for every matched pair (A tile ta, B tile tb) of C tile
// (I, J), in ascending inner tile index K, uniformly across the 16 lanes of
// the calling group (g = group in the wave, r = lane in the group).  The
// keys of the shorter list are binary-searched in the longer one, 16 per
// round; a 64-bit ballot keeps the matches in key order.  Short lists are
// first copied to the group's LDS slots (sa/sb) so the searches hit LDS.
template <typename Body>
__device__ __forceinline__ void for_each_pair(const TileOperands &op, int I, int J, int g, int r, int *sa, int *sb,
                                              Body body) {
    const int a0 = op.Aptr[I], la = op.Aptr[I + 1] - a0;
    const int b0 = op.Bcscptr[J], lb = op.Bcscptr[J + 1] - b0;
    if (la == 0 || lb == 0) return;
    const bool a_short = la <= lb;
    const int ls = a_short ? la : lb, ll = a_short ? lb : la;
    const int *S = a_short ? op.Acol + a0 : op.Bcscrow + b0;
    const int *L = a_short ? op.Bcscrow + b0 : op.Acol + a0;
    const bool staged = la <= PAIR_LDS_MAX && lb <= PAIR_LDS_MAX;
    if (staged) {
        for (int x = r; x < ls; x += 16) sa[x] = S[x];
        for (int x = r; x < ll; x += 16) sb[x] = L[x];
        tsg_wave_sync();
        S = sa;
        L = sb;
    }
    const int src_base = g * 16;
    for (int base = 0; base < ls; base += 16) {
        const int idx = base + r;
        int pos = -1;
        if (idx < ls) pos = find_sorted(L, ll, S[idx]);
        const unsigned long long bal = __ballot(pos >= 0);
        unsigned m16 = (unsigned)(bal >> src_base) & 0xffffu;
        while (m16) {
            const int j = __ffs(m16) - 1;
            m16 &= m16 - 1;
            const int sidx = base + j;
            const int lpos = __shfl(pos, src_base + j, 64);
            const int apos = a_short ? sidx : lpos;
            const int bpos = a_short ? lpos : sidx;
            body(a0 + apos, op.Bcscid[b0 + bpos]);
        }
    }
    if (staged) tsg_wave_sync();
}

// ------------------------------------------------------------------ segments
// A segment is a range [s0, s1) of the C tiles of one tile-row I, stored
// as int4 {I, s0, s1, 0}.  Long tile-rows are cut into segments so each
// fits a block's LDS; a segment's C tiles cover the tile columns
// [Ccol[c0+s0], Ccol[c0+s1-1]], and only the B tiles inside that column
// range meet them.  MODE 0 cuts by tile count; MODE 1 also keeps every
// segment within max_nnz C nonzeros (tile nnz <= 256 <= max_nnz).  Only
// tile-rows with rowcls == want (want < 0: all) are cut.
template <int MODE>
__device__ __forceinline__ int seg_walk(const int *__restrict__ Cnnz, int c0, int u, int max_tiles, int max_nnz,
                                        int4 *out, int I, int kind = 0) {
    if (MODE == 0) {
        const int n = (u + max_tiles - 1) / max_tiles;
        if (out)
            for (int j = 0; j < n; j++) out[j] = make_int4(I, j * max_tiles, min(u, (j + 1) * max_tiles), kind);
        return n;
    }
    // MODE 1: tile s belongs to segment floor(p(s) / H), p(s) = C nonzeros of
    // the row before tile s and H = max_nnz - 255; a tile holds at most 256
    // nonzeros, so ids grow by at most one per tile, every id up to the last
    // is present, and a segment holds at most H + 255 = max_nnz nonzeros.
    // Count in O(1), each boundary by one binary search (no tile walk).
    const int H = max_nnz - 255;
    const int base = Cnnz[c0];
    const int n = (Cnnz[c0 + u - 1] - base) / H + 1;
    if (out) {
        int s0 = 0;
        for (int j = 0; j < n; j++) {
            int s1 = u;
            if (j + 1 < n) {  // first tile with p >= (j+1) H
                const int key = base + (j + 1) * H;
                int lo = s0, hi = u - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (Cnnz[c0 + mid] < key) lo = mid + 1; else hi = mid;
                }
                s1 = lo;
            }
            out[j] = make_int4(I, s0, s1, j);  // .w: index of the segment in its tile-row
            s0 = s1;
        }
    }
    return n;
}

template <int MODE>
__global__ void __launch_bounds__(256)
k_seg(const int *__restrict__ Cptr, const int *__restrict__ Cnnz, int tilem, const unsigned char *__restrict__ rowcls,
      int want, int max_tiles, int max_nnz, int *__restrict__ cnt_or_off, int4 *__restrict__ items) {
    const int I = blockIdx.x * 256 + threadIdx.x;
    if (I >= tilem) return;
    const int c0 = Cptr[I], u = Cptr[I + 1] - c0;
    // MODE 0 carries rowcls (the step-2 kernel kind) into each item's .w
    const bool take = u > 0 && (want < 0 || rowcls[I] == want);
    const int kind = (MODE == 0 && rowcls) ? rowcls[I] : 0;
    if (!items) {
        cnt_or_off[I] = take ? seg_walk<MODE>(Cnnz, c0, u, max_tiles, max_nnz, nullptr, I, kind) : 0;
    } else if (take) {
        seg_walk<MODE>(Cnnz, c0, u, max_tiles, max_nnz, items + cnt_or_off[I], I, kind);
    }
}

// Step-2 kernel kind of every tile-row: 1 = element-wise (B has its row
// index and A's tile-row is hypersparse: at most 4 entries per tile on
// average, or every row when all_elem), 0 = tile pairs.
__global__ void __launch_bounds__(256)
k_step2_kind(const int *__restrict__ Aptr, const int *__restrict__ Annz, int tilem, int all_elem,
             unsigned char *__restrict__ kind) {
    const int I = blockIdx.x * 256 + threadIdx.x;
    if (I >= tilem) return;
    const int a0 = Aptr[I], a1 = Aptr[I + 1];
    const long long e = (long long)Annz[a1] - Annz[a0];
    kind[I] = (unsigned char)((all_elem || e <= 4LL * (a1 - a0)) ? 1 : 0);
}

// Upper bound on the segments of a C with tilem tile-rows and numtile
// tiles: one per tile-row, one per max_tiles tiles closed by count, and
// (MODE 1) one per (max_nnz - 255) nonzeros closed by size, with at most
// 256 nonzeros per tile.
static long long seg_bound(long long tilem, long long numtile, int max_tiles, int max_nnz, int mode) {
    long long b = tilem + numtile / max_tiles + 1;
    if (mode == 1) b += 256 * numtile / (max_nnz - 255) + 1;
    return b;
}

// First index in [lo, hi) with list[index] >= key (list ascending).
__device__ __forceinline__ int lower_bound_i(const int *__restrict__ list, int lo, int hi, int key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// MODE 1 segments, one thread per segment (a hub tile-row has thousands, each
// found by a binary search; one thread per row left the whole launch
// waiting on the longest row): segment j's row is the last row whose first
// segment is at or before j (off[] is the scanned per-row count, off[tilem]
// the total), and its bounds are the same tiles seg_walk finds.
__global__ void __launch_bounds__(256)
k_seg1_fill(const int *__restrict__ Cptr, const int *__restrict__ Cnnz, int tilem, const int *__restrict__ off,
            int max_nnz, int4 *__restrict__ items) {
    const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
    if (j >= off[tilem]) return;
    int lo = 0, hi = tilem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const int I = lo, jj = (int)(j - off[I]), n = off[I + 1] - off[I];
    const int c0 = Cptr[I], u = Cptr[I + 1] - c0;
    const int H = max_nnz - 255, base = Cnnz[c0];
    const int s0 = jj == 0 ? 0 : lower_bound_i(Cnnz, c0, c0 + u - 1, base + jj * H) - c0;
    const int s1 = jj + 1 == n ? u : lower_bound_i(Cnnz, c0, c0 + u - 1, base + (jj + 1) * H) - c0;
    items[j] = make_int4(I, s0, s1, jj);
}

// ------------------------------------------------------------------ step 2
// Row-wise (Gustavson) bitmask symbolic: one 256-thread block per segment
// of a C tile-row.  Every (A tile (I,K), B tile (K,J)) pair of the row
// with J in the segment's column range -- the pairs step 1 enumerated --
// is visited exactly once, flattened over the threads; no intersection
// search is needed.  Row r of the pair's product pattern, OR over k in
// Amask[r] of Bmask[k], is OR-ed into C tile J's row mask held in LDS (OR
// is order-free, so LDS atomics are exact).  A tiles with a single entry
// (r, k) contribute B row k to C row r alone; their pairs are enumerated
// before the others' so that a wave's lanes take one kind of pair.
#define G2_THREADS 256
#define G2_UMAX 512    // C tiles of one segment held in LDS
#define G2_PCHUNK 256  // A tiles per enumeration round
#define G2_MAP_MAX 16384  // widest tile-row span with a direct column map (bytes of LDS)
#define G2_MSTRIDE 9      // LDS words per C tile mask set / per thread B mask copy (8 used)

// Most C tiles of a tile-row (capped at the step-2 segment size), the
// widest tile-column span of a tile-row, and the widest span of at most
// G2_MAP_MAX: out[0], out[1], out[2] (atomic maxima), from the step-1 count
// pass's per-row tile counts and spans.
// A few large blocks stride over the rows, so that the atomics (one cache
// line, serialised) are few.
#define ROW_STATS_BLOCKS 64
__global__ void __launch_bounds__(1024)
k_row_stats(const int *__restrict__ cnt, const int *__restrict__ span, int tilem, int *__restrict__ out) {
    __shared__ int s_r[16][3];
    int nu = 0, sp = 0, ms = 0;
    for (int I = blockIdx.x * 1024 + threadIdx.x; I < tilem; I += gridDim.x * 1024) {
        if (cnt[I] > 0) {
            const int s1 = span[I];
            nu = max(nu, min(cnt[I], G2_UMAX));
            sp = max(sp, s1);
            if (s1 <= G2_MAP_MAX) ms = max(ms, s1);
        }
    }
    nu = tsg_wave_max(nu);
    sp = tsg_wave_max(sp);
    ms = tsg_wave_max(ms);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_r[wave][0] = nu;
        s_r[wave][1] = sp;
        s_r[wave][2] = ms;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int v = s_r[0][threadIdx.x];
        for (int w = 1; w < 16; w++) v = max(v, s_r[w][threadIdx.x]);
        atomicMax(&out[threadIdx.x], v);
    }
}

static_assert(G2_PCHUNK == G2_THREADS, "one A tile of a chunk per thread");
__global__ void __launch_bounds__(G2_THREADS)
k_step2_rows(TileOperands op, const int4 *__restrict__ segs, const int *__restrict__ nseg, const int *__restrict__ Cptr,
             const int *__restrict__ Ccol, unsigned short *__restrict__ Cmask, unsigned char *__restrict__ Cptr16,
             int *__restrict__ Cnnz, int umax, int mapbytes) {
    // dynamic LDS, sized by the largest segment: the segment's C tile masks
    // (two 16-bit rows per word, even row high; G2_MSTRIDE words per tile so
    // that the lanes' tiles fall in distinct banks), its tile columns, and
    // when every segment has at most 256 tiles, mapbytes bytes for a direct
    // map tile column -> slot, used by the segments spanning at most
    // mapbytes tile columns (the others search the tile columns)
    extern __shared__ __align__(16) unsigned char g2_smem[];
    unsigned *s_m = reinterpret_cast<unsigned *>(g2_smem);
    int *s_ccol = reinterpret_cast<int *>(g2_smem + 4 * G2_MSTRIDE * (size_t)umax);
    unsigned char *s_map = g2_smem + (4 * G2_MSTRIDE + 4) * (size_t)umax;
    // per enumeration position (single-entry A tiles first): first pair,
    // first B tile, A tile (offset in the chunk), single entry (row |
    // column << 4, or 255)
    __shared__ int s_pre[G2_PCHUNK];
    __shared__ int s_b0[G2_PCHUNK];
    __shared__ unsigned char s_pid[G2_PCHUNK];
    __shared__ unsigned char s_rk[G2_PCHUNK];
    __shared__ int s_wsg[G2_THREADS / 64];
    __shared__ int s_tot;
    // B masks of each thread's pair, word-major (word i = rows 2i (low half)
    // and 2i+1), so that a wave's lanes read and write consecutive banks
    __shared__ unsigned s_bst[8][G2_THREADS];
    const int tid = threadIdx.x;
    // the grid is an upper bound: the first *nseg blocks (spread evenly over
    // the XCDs) take the segments, each XCD a contiguous range
    const int ns = *nseg;
    if ((int)blockIdx.x >= ns) return;  // block-uniform
    const int w = (int)tsg_xcd_block(blockIdx.x, (unsigned)ns);
    const int4 sg = segs[w];
    if (sg.w) return;  // element-wise segment (k_step2_elem)
    const int I = sg.x;
    const int c0 = Cptr[I] + sg.y, u = sg.z - sg.y;
    const bool whole = sg.y == 0 && sg.z == Cptr[I + 1] - Cptr[I];
    const int jlo = Ccol[c0], jhi = Ccol[c0 + u - 1];
    const bool usemap = jhi - jlo < mapbytes;
    for (int i = tid; i < u; i += G2_THREADS) {
        const int J = Ccol[c0 + i];
        s_ccol[i] = J;
        if (usemap) s_map[J - jlo] = (unsigned char)i;
    }
    for (int i = tid; i < u * G2_MSTRIDE; i += G2_THREADS) s_m[i] = 0u;
    const int a0 = op.Aptr[I], a1 = op.Aptr[I + 1];
    __shared__ int s_wsum[G2_THREADS / 64];
    for (int pc = a0; pc < a1; pc += G2_PCHUNK) {
        const int np = min(G2_PCHUNK, a1 - pc);
        int len = 0, b0 = 0, rk = 255;
        if (tid < np) {
            const uint4 *am4 = reinterpret_cast<const uint4 *>(op.Amask + (long long)(pc + tid) * 16);
            const uint4 aA = am4[0], aB = am4[1];
            const unsigned aw[8] = {aA.x, aA.y, aA.z, aA.w, aB.x, aB.y, aB.z, aB.w};
            int nz = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                nz += __popc(aw[i]);
                if (aw[i]) {
                    const int bit = 31 - __clz(aw[i]);  // row 2i in the low half, 2i+1 in the high
                    rk = (2 * i + (bit >> 4)) | ((15 - (bit & 15)) << 4);
                }
            }
            if (nz != 1) rk = 255;
            const int K = op.Acol[pc + tid];
            b0 = op.Bptr[K];
            int b1 = op.Bptr[K + 1];
            if (!whole) {  // B tiles of row K inside the segment's columns
                b0 = lower_bound_i(op.Bcol, b0, b1, jlo);
                b1 = lower_bound_i(op.Bcol, b0, b1, jhi + 1);
            }
            len = b1 - b0;
        }
        // enumeration positions: single-entry A tiles first, in thread order
        {
            const bool sgl = tid < np && rk != 255, dns = tid < np && rk == 255;
            const unsigned long long sb = __ballot(sgl), db = __ballot(dns);
            const unsigned long long lt = (1ull << (tid & 63)) - 1ull;
            if ((tid & 63) == 0) {
                s_wsg[tid >> 6] = __popcll(sb);
                s_wsum[tid >> 6] = __popcll(db);
            }
            __syncthreads();
            int ns = 0, sbef = 0, dbef = 0;
            for (int w2 = 0; w2 < G2_THREADS / 64; w2++) {
                ns += s_wsg[w2];
                if (w2 < (tid >> 6)) {
                    sbef += s_wsg[w2];
                    dbef += s_wsum[w2];
                }
            }
            const int pos = sgl ? sbef + __popcll(sb & lt) : ns + dbef + __popcll(db & lt);
            if (tid < np) {
                s_pre[pos] = len;
                s_b0[pos] = b0;
                s_pid[pos] = (unsigned char)tid;
                s_rk[pos] = (unsigned char)rk;
            }
            __syncthreads();
            len = tid < np ? s_pre[tid] : 0;
        }
        // block-wide exclusive scan of len in position order
        const int incl = tsg_wave_incl_scan(len);
        if ((tid & 63) == 63) s_wsum[tid >> 6] = incl;
        __syncthreads();
        int woff = 0;
        for (int w2 = 0; w2 < (tid >> 6); w2++) woff += s_wsum[w2];
        s_pre[tid] = woff + incl - len;
        if (tid == G2_THREADS - 1) s_tot = woff + incl;
        __syncthreads();
        const int total = s_tot;
        // two pairs per round, so that both pairs' B tile loads are in flight
#pragma unroll 2
        for (int x = tid; x < total; x += G2_THREADS) {
            int p = 0;
#pragma unroll
            for (int st = G2_PCHUNK / 2; st > 0; st >>= 1)
                if (p + st < G2_PCHUNK && s_pre[p + st] <= x) p += st;
            const int q = s_b0[p] + (x - s_pre[p]);
            const int J = op.Bcol[q];  // J is in the segment's step-1 list
            int slot;
            if (usemap) {
                slot = s_map[J - jlo];
            } else {
                int lo = 0, hi = u - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_ccol[mid] < J) lo = mid + 1; else hi = mid;
                }
                slot = lo;
            }
            const int rk1 = s_rk[p];
            if (rk1 != 255) {  // single entry (r, k): B row k into C row r
                const int r = rk1 & 15;
                const unsigned m = op.Bmask[(long long)q * 16 + (rk1 >> 4)];
                atomicOr(&s_m[slot * G2_MSTRIDE + (r >> 1)], m << ((r & 1) ? 0 : 16));
                continue;
            }
            const uint4 *bm4 = reinterpret_cast<const uint4 *>(op.Bmask + (long long)q * 16);
            const uint4 bA = bm4[0], bB = bm4[1];
            s_bst[0][tid] = bA.x;
            s_bst[1][tid] = bA.y;
            s_bst[2][tid] = bA.z;
            s_bst[3][tid] = bA.w;
            s_bst[4][tid] = bB.x;
            s_bst[5][tid] = bB.y;
            s_bst[6][tid] = bB.z;
            s_bst[7][tid] = bB.w;
            const uint4 *am4 = reinterpret_cast<const uint4 *>(op.Amask + (long long)(pc + s_pid[p]) * 16);
            const uint4 aA = am4[0], aB = am4[1];
            const unsigned aw[8] = {aA.x, aA.y, aA.z, aA.w, aB.x, aB.y, aB.z, aB.w};
#pragma unroll
            for (int r = 0; r < 16; r++) {
                unsigned am = (aw[r >> 1] >> ((r & 1) * 16)) & 0xffffu;  // little-endian: row 2w low half
                if (!am) continue;
                unsigned m = 0;
                while (am) {
                    const int b = 31 - __clz(am);
                    am &= ~(1u << b);
                    const int k = 15 - b;
                    m |= (s_bst[k >> 1][tid] >> ((k & 1) * 16)) & 0xffffu;
                }
                atomicOr(&s_m[slot * G2_MSTRIDE + (r >> 1)], m << ((r & 1) ? 0 : 16));
            }
        }
        __syncthreads();
    }
    // per C tile: popcounts, u8 row offsets, tile nnz
    for (int s = tid; s < u; s += G2_THREADS) {
        unsigned w8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) w8[k] = s_m[s * G2_MSTRIDE + k];
        unsigned mk[8], pt[4];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const unsigned hi = w8[k] >> 16, lo = w8[k] & 0xffffu;  // rows 2k, 2k+1
            mk[k] = hi | (lo << 16);                               // little-endian u16 pair
        }
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
            unsigned o = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int r = q4 * 4 + b;
                const unsigned row = (r & 1) ? (w8[r >> 1] & 0xffffu) : (w8[r >> 1] >> 16);
                o |= (unsigned)sum << (8 * b);
                sum += __popc(row);
            }
            pt[q4] = o;
        }
        const long long t = c0 + s;
        uint4 *cm4 = reinterpret_cast<uint4 *>(Cmask + t * 16);
        cm4[0] = make_uint4(mk[0], mk[1], mk[2], mk[3]);
        cm4[1] = make_uint4(mk[4], mk[5], mk[6], mk[7]);
        *reinterpret_cast<uint4 *>(Cptr16 + t * 16) = make_uint4(pt[0], pt[1], pt[2], pt[3]);
        Cnnz[t] = sum;
    }
}

// Element-wise bitmask symbolic for hypersparse tile-rows (B carries its
// row index): one 256-thread block per segment; each thread takes whole A
// tiles of the tile-row and, for each of their entries (r, k), ORs the
// columns of B row k that fall in the segment's range into the row-r mask
// of their C tile (slot found through an LDS hash of the segment's tile
// columns).  Work is one step per intermediate product instead of one per
// (A tile, B tile) pair; OR is order-free, so LDS atomics are exact.
#define G2E_HASH 1024  // >= 2 * G2_UMAX
static_assert(G2_UMAX <= 512, "C tile index in 9 bits of a step-2 hash slot");

__device__ __forceinline__ unsigned col_hash10(int J) { return ((unsigned)J * 2654435761u) >> 22; }

__global__ void __launch_bounds__(G2_THREADS)
k_step2_elem(TileOperands op, const int *__restrict__ Browp, const int *__restrict__ Bcolx,
             const int4 *__restrict__ segs, const int *__restrict__ nseg, const int *__restrict__ Cptr,
             const int *__restrict__ Ccol, unsigned short *__restrict__ Cmask, unsigned char *__restrict__ Cptr16,
             int *__restrict__ Cnnz) {
    __shared__ unsigned s_m[G2_UMAX * 8];  // two 16-bit rows per word, even row high
    // hash of the segment's tile columns: slot (J - jlo + 1) << 9 | C tile
    // index (J - jlo + 1 < 2^23 as B has at most 2^27 - 32 columns when this
    // kernel runs, index < G2_UMAX = 512); 0 = empty
    __shared__ unsigned s_hkey[G2E_HASH];
    const int tid = threadIdx.x;
    const int ns = *nseg;
    if ((int)blockIdx.x >= ns) return;  // block-uniform
    const int w = (int)tsg_xcd_block(blockIdx.x, (unsigned)ns);
    const int4 sg = segs[w];
    if (!sg.w) return;  // tile-pair segment (k_step2_rows)
    const int I = sg.x;
    const int c0 = Cptr[I] + sg.y, u = sg.z - sg.y;
    for (int i = tid; i < u * 8; i += G2_THREADS) s_m[i] = 0u;
    for (int i = tid; i < G2E_HASH; i += G2_THREADS) s_hkey[i] = 0u;
    __syncthreads();
    const int jlo = Ccol[c0];
    for (int i = tid; i < u; i += G2_THREADS) {
        const int J = Ccol[c0 + i];
        unsigned h = col_hash10(J);
        const unsigned ent = (unsigned)(J - jlo + 1) << 9 | (unsigned)i;
        while (atomicCAS(&s_hkey[h], 0u, ent) != 0u) h = (h + 1) & (G2E_HASH - 1);  // u <= G2E_HASH / 2
    }
    __syncthreads();
    const int col_lo = Ccol[c0] * 16, col_hi = Ccol[c0 + u - 1] * 16 + 16;
    const int a0 = op.Aptr[I], a1 = op.Aptr[I + 1];
    __shared__ int s_tex[G2_THREADS + 1];  // entries before each tile of the chunk
    for (int tc = a0; tc < a1; tc += G2_THREADS) {
        // one chunk of A tiles: their entries numbered block-wide, one
        // entry per thread per step (balanced whatever the tile densities)
        const int t = tc + tid;
        int cnt = 0;
        if (t < a1) {
            const uint4 *am4 = reinterpret_cast<const uint4 *>(op.Amask + (long long)t * 16);
            const uint4 aA = am4[0], aB = am4[1];
            cnt = __popc(aA.x) + __popc(aA.y) + __popc(aA.z) + __popc(aA.w) + __popc(aB.x) + __popc(aB.y) +
                  __popc(aB.z) + __popc(aB.w);
        }
        __shared__ int s_w[G2_THREADS / 64];
        const int incl = tsg_wave_incl_scan(cnt);
        if ((tid & 63) == 63) s_w[tid >> 6] = incl;
        __syncthreads();
        int woff = 0;
        for (int q = 0; q < (tid >> 6); q++) woff += s_w[q];
        s_tex[tid] = woff + incl - cnt;
        if (tid == G2_THREADS - 1) s_tex[G2_THREADS] = woff + incl;
        __syncthreads();
        const int total = s_tex[G2_THREADS];
        for (int x = tid; x < total; x += G2_THREADS) {
            int ow = 0;  // tile of entry x: last with s_tex <= x
#pragma unroll
            for (int st = G2_THREADS / 2; st > 0; st >>= 1)
                if (ow + st < G2_THREADS && s_tex[ow + st] <= x) ow += st;
            const int tt = tc + ow;
            int j = x - s_tex[ow];
            // the tile's 16 row masks in one load; the row holding its j-th entry
            const uint4 *am4 = reinterpret_cast<const uint4 *>(op.Amask + (long long)tt * 16);
            const uint4 mA = am4[0], mB = am4[1];
            const unsigned mw[8] = {mA.x, mA.y, mA.z, mA.w, mB.x, mB.y, mB.z, mB.w};
            int r = 0;
            unsigned am = mw[0] & 0xffffu;
#pragma unroll
            for (int rr = 0; rr < 16; rr++) {
                const unsigned mr = (mw[rr >> 1] >> ((rr & 1) * 16)) & 0xffffu;
                if (rr == r && r < 15 && j >= __popc(mr)) {
                    j -= __popc(mr);
                    r++;
                }
            }
            am = (mw[r >> 1] >> ((r & 1) * 16)) & 0xffffu;
            for (int q = 0; q < j; q++) am &= ~(1u << (31 - __clz(am)));
            const int k = op.Acol[tt] * 16 + 15 - (31 - __clz(am));
            const int e1 = Browp[k + 1];
            bool more = true;
            // B row k's entries in the segment's column range, 4 loads at a time
            // the tile-row's first segment starts at B row k's first entry
            const int eb = sg.y == 0 ? Browp[k] : lower_bound_i(Bcolx, Browp[k], e1, col_lo);
            for (int e = eb; more && e < e1; e += 4) {
                int cols[4];
#pragma unroll
                for (int v = 0; v < 4; v++) cols[v] = e + v < e1 ? Bcolx[e + v] : 0x7fffffff;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int col = cols[v];
                    if (col >= col_hi) more = false;
                    if (!more) continue;
                    const int J = col >> 4;
                    unsigned h = col_hash10(J);  // J is in the segment's tile list
                    const unsigned key = (unsigned)(J - jlo + 1);
                    int probe = 0;
                    unsigned ent;
                    while (((ent = s_hkey[h]) >> 9) != key && ++probe < G2E_HASH) h = (h + 1) & (G2E_HASH - 1);
                    if (probe == G2E_HASH) continue;  // not reached for a consistent C structure
                    const unsigned bit = 1u << (15 - (col & 15));
                    atomicOr(&s_m[(int)(ent & 511u) * 8 + (r >> 1)], bit << ((r & 1) ? 0 : 16));
                }
            }
        }
        __syncthreads();  // s_tex is rewritten by the next chunk
    }
    __syncthreads();
    for (int sl = tid; sl < u; sl += G2_THREADS) {
        unsigned w8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) w8[k] = s_m[sl * 8 + k];
        unsigned mk[8], pt[4];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) mk[k] = (w8[k] >> 16) | ((w8[k] & 0xffffu) << 16);
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
            unsigned o = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int r = q4 * 4 + b;
                const unsigned row = (r & 1) ? (w8[r >> 1] & 0xffffu) : (w8[r >> 1] >> 16);
                o |= (unsigned)sum << (8 * b);
                sum += __popc(row);
            }
            pt[q4] = o;
        }
        const long long t = c0 + sl;
        uint4 *cm4 = reinterpret_cast<uint4 *>(Cmask + t * 16);
        cm4[0] = make_uint4(mk[0], mk[1], mk[2], mk[3]);
        cm4[1] = make_uint4(mk[4], mk[5], mk[6], mk[7]);
        *reinterpret_cast<uint4 *>(Cptr16 + t * 16) = make_uint4(pt[0], pt[1], pt[2], pt[3]);
        Cnnz[t] = sum;
    }
}

// Tile-parallel step 2 for tile-rows with more than G2_UMAX C tiles: one C
// tile per 16-lane group, pairs from the sorted-list intersection.
// Tile-parallel grids cover 16 work-items per C tile, which exceeds the
// 2^32 work-items of a launch beyond 2^28 tiles: such kernels run at most
// TP_MAX_BLOCKS blocks, each looping over virtual blocks.
#define TP_MAX_BLOCKS (1u << 20)

static inline unsigned tp_grid(long long vblocks) {
    return (unsigned)std::min<long long>(std::max<long long>(vblocks, 1), TP_MAX_BLOCKS);
}

__device__ __forceinline__ void step2_tile_body(const TileOperands &op, const int *__restrict__ Cptr,
                                                const int *__restrict__ Crow, const int *__restrict__ Ccol,
                                                int numtileC, unsigned short *__restrict__ Cmask,
                                                unsigned char *__restrict__ Cptr16, int *__restrict__ Cnnz,
                                                int all_rows, long long vblk, unsigned short *s_bm,
                                                int (*s_la)[PAIR_LDS_MAX], int (*s_lb)[PAIR_LDS_MAX]);

__global__ void __launch_bounds__(64 * S23_WAVES)
k_step2_tiles(TileOperands op, const int *__restrict__ Cptr, const int *__restrict__ Crow,
              const int *__restrict__ Ccol, int numtileC, unsigned short *__restrict__ Cmask,
              unsigned char *__restrict__ Cptr16, int *__restrict__ Cnnz, int all_rows) {
    __shared__ unsigned short s_bm[S23_WAVES][64];
    __shared__ int s_la[S23_WAVES][TILES_PER_WAVE][PAIR_LDS_MAX];
    __shared__ int s_lb[S23_WAVES][TILES_PER_WAVE][PAIR_LDS_MAX];
    const int wave = threadIdx.x >> 6;
    const long long nvb = ((long long)numtileC + S23_WAVES * TILES_PER_WAVE - 1) / (S23_WAVES * TILES_PER_WAVE);
    for (long long vb = tsg_xcd_block(blockIdx.x, gridDim.x); vb < nvb; vb += gridDim.x)
        step2_tile_body(op, Cptr, Crow, Ccol, numtileC, Cmask, Cptr16, Cnnz, all_rows, vb, s_bm[wave], s_la[wave],
                        s_lb[wave]);
}

__device__ __forceinline__ void step2_tile_body(const TileOperands &op, const int *__restrict__ Cptr,
                                                const int *__restrict__ Crow, const int *__restrict__ Ccol,
                                                int numtileC, unsigned short *__restrict__ Cmask,
                                                unsigned char *__restrict__ Cptr16, int *__restrict__ Cnnz,
                                                int all_rows, long long vblk, unsigned short *s_bmw,
                                                int (*s_law)[PAIR_LDS_MAX], int (*s_lbw)[PAIR_LDS_MAX]) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const long long t = (vblk * S23_WAVES + wave) * TILES_PER_WAVE + g;
    if (t >= numtileC) return;  // whole groups exit together
    const int I = Crow[t], J = Ccol[t];
    if (!all_rows && Cptr[I + 1] - Cptr[I] <= G2_UMAX) return;  // done by k_step2_rows
    unsigned short *bmg = &s_bmw[g * 16];
    unsigned orow = 0;
    for_each_pair(op, I, J, g, r, s_law[g], s_lbw[g], [&](int ta, int tb) {
        bmg[r] = op.Bmask[(long long)tb * 16 + r];
        tsg_wave_sync();
        unsigned am = op.Amask[(long long)ta * 16 + r];
        while (am) {
            const int b = 31 - __clz(am);
            am &= ~(1u << b);
            orow |= bmg[15 - b];
        }
        tsg_wave_sync();
    });
    const int cnt = __popc(orow);
    const int incl = tsg_group16_incl_scan(cnt);
    const int tot = __shfl(incl, g * 16 + 15, 64);
    Cmask[t * 16 + r] = (unsigned short)orow;
    Cptr16[t * 16 + r] = (unsigned char)(incl - cnt);
    if (r == 0) Cnnz[t] = tot;
}

// Tile counts per accumulator class (statistics only): one atomic per block.
__global__ void __launch_bounds__(256)
k_bins(const int *__restrict__ cnt, long long n, int dense_min, unsigned long long *__restrict__ bins) {
    __shared__ unsigned s[3];
    if (threadIdx.x < 3) s[threadIdx.x] = 0;
    __syncthreads();
    unsigned c0 = 0, c1 = 0, c2 = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int v = cnt[i];
        c0 += v == 0;
        c1 += v > 0 && v < dense_min;
        c2 += v >= dense_min;
    }
    c0 = tsg_wave_sum(c0);
    c1 = tsg_wave_sum(c1);
    c2 = tsg_wave_sum(c2);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&s[0], c0);
        atomicAdd(&s[1], c1);
        atomicAdd(&s[2], c2);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd(&bins[threadIdx.x], (unsigned long long)s[threadIdx.x]);
}

__global__ void __launch_bounds__(256)
k_max_rowlen(const int *__restrict__ ptr, int n, int *__restrict__ out) {
    int m = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = max(m, ptr[i + 1] - ptr[i]);
    m = tsg_wave_max(m);
    __shared__ int s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(s[0], s[1]), max(s[2], s[3])));
}

// ------------------------------------------------------------------ step 3
// Tile-rows whose C tiles and nonzeros fit the row-wise kernel's LDS
// accumulator go through k_step3_rows; the others through the
// tile-parallel k_step3 below (any size, no LDS limit on the row).
#define G3_THREADS 256
// Caps of the row-wise kernel (the LDS layout is sized at launch from the
// actual maxima of the rows that fit, at most ~60 KB per block).
#define G3_UMAX 256     // C tiles per tile-row (u8 slot ids)
#define G3_NNZMAX 3072  // C nonzeros per tile-row (LDS accumulator)
#define G3_QB 32        // B tiles per B tile-row (staging thread roles)
#define G3_BNNZ 512     // B values per B tile-row (two per thread)


#define ACC_STRIDE 17  // doubles per lane row (padding against bank conflicts)

__device__ __forceinline__ void step3_tile_body(const TileOperands &op, const int *__restrict__ Crow,
                                                const int *__restrict__ Ccol, int numtileC,
                                                const unsigned short *__restrict__ Cmask,
                                                const unsigned char *__restrict__ Cptr16,
                                                const int *__restrict__ Cnnz, unsigned char *__restrict__ Ccol8,
                                                double *__restrict__ Cval, int dense_min, int all_rows,
                                                const unsigned char *__restrict__ rowcls, long long vblk,
                                                double *s_accw, unsigned short *s_bmw, unsigned char *s_bpw,
                                                int (*s_law)[PAIR_LDS_MAX], int (*s_lbw)[PAIR_LDS_MAX]);

__global__ void __launch_bounds__(64 * S23_WAVES)
k_step3(TileOperands op, const int *__restrict__ Cptr, const int *__restrict__ Crow, const int *__restrict__ Ccol,
        int numtileC, const unsigned short *__restrict__ Cmask, const unsigned char *__restrict__ Cptr16,
        const int *__restrict__ Cnnz, unsigned char *__restrict__ Ccol8, double *__restrict__ Cval,
        int dense_min, int all_rows, const unsigned char *__restrict__ rowcls) {
    __shared__ double s_acc[S23_WAVES][64 * ACC_STRIDE];
    __shared__ unsigned short s_bm[S23_WAVES][64];
    __shared__ unsigned char s_bp[S23_WAVES][64];
    __shared__ int s_la[S23_WAVES][TILES_PER_WAVE][PAIR_LDS_MAX];
    __shared__ int s_lb[S23_WAVES][TILES_PER_WAVE][PAIR_LDS_MAX];
    const int wave = threadIdx.x >> 6;
    const long long nvb = ((long long)numtileC + S23_WAVES * TILES_PER_WAVE - 1) / (S23_WAVES * TILES_PER_WAVE);
    for (long long vb = tsg_xcd_block(blockIdx.x, gridDim.x); vb < nvb; vb += gridDim.x)
        step3_tile_body(op, Crow, Ccol, numtileC, Cmask, Cptr16, Cnnz, Ccol8, Cval, dense_min, all_rows, rowcls, vb,
                        s_acc[wave], s_bm[wave], s_bp[wave], s_la[wave], s_lb[wave]);
}

__device__ __forceinline__ void step3_tile_body(const TileOperands &op, const int *__restrict__ Crow,
                                                const int *__restrict__ Ccol, int numtileC,
                                                const unsigned short *__restrict__ Cmask,
                                                const unsigned char *__restrict__ Cptr16,
                                                const int *__restrict__ Cnnz, unsigned char *__restrict__ Ccol8,
                                                double *__restrict__ Cval, int dense_min, int all_rows,
                                                const unsigned char *__restrict__ rowcls, long long vblk,
                                                double *s_accw, unsigned short *s_bmw, unsigned char *s_bpw,
                                                int (*s_law)[PAIR_LDS_MAX], int (*s_lbw)[PAIR_LDS_MAX]) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const long long t = (vblk * S23_WAVES + wave) * TILES_PER_WAVE + g;
    if (t >= numtileC) return;
    const int coff = Cnnz[t], cn = Cnnz[t + 1] - coff;
    if (cn == 0) return;  // structurally empty tile (step-1 phantom)
    if (!all_rows && rowcls[Crow[t]] == 1) return;  // done by k_step3_rows
    const unsigned cm = Cmask[t * 16 + r];
    const int cp = Cptr16[t * 16 + r];
    const int rowlen = __popc(cm);
    const bool dense = cn >= dense_min;
    double *accg = &s_accw[g * 16 * ACC_STRIDE];  // 272 doubles per group
    double *accr = accg + r * ACC_STRIDE;          // dense: lane-private row
    if (dense) {
#pragma unroll
        for (int c = 0; c < 16; c++) accr[c] = 0.0;
    } else {
        for (int j = 0; j < rowlen; j++) accg[cp + j] = 0.0;
    }
    unsigned short *bmg = &s_bmw[g * 16];
    unsigned char *bpg = &s_bpw[g * 16];
    const int I = Crow[t], J = Ccol[t];
    for_each_pair(op, I, J, g, r, s_law[g], s_lbw[g], [&](int ta, int tb) {
        bmg[r] = op.Bmask[(long long)tb * 16 + r];
        bpg[r] = op.Bptr16[(long long)tb * 16 + r];
        tsg_wave_sync();
        const double *bval = op.Bval + op.Bnnz[tb];
        unsigned am = op.Amask[(long long)ta * 16 + r];
        const double *aval = op.Aval + op.Annz[ta] + op.Aptr16[(long long)ta * 16 + r];
        int ia = 0;
        while (am) {
            const int b = 31 - __clz(am);  // highest bit = smallest column k
            am &= ~(1u << b);
            const int k = 15 - b;
            const double a = aval[ia++];
            unsigned bm = bmg[k];
            const double *brow = bval + bpg[k];
            int jb = 0;
            while (bm) {
                const int bb = 31 - __clz(bm);
                bm &= ~(1u << bb);
                const double bv = brow[jb++];
                double *slot = dense ? &accr[15 - bb] : &accg[cp + __popc(cm >> (bb + 1))];
                *slot = __builtin_fma(a, bv, *slot);
            }
        }
        tsg_wave_sync();
    });
    // write out row r of the tile, ascending columns
    double *outv = Cval + coff + cp;
    unsigned char *outc = Ccol8 + coff + cp;
    unsigned x = cm;
    for (int j = 0; j < rowlen; j++) {
        const int bb = 31 - __clz(x);
        x &= ~(1u << bb);
        const int c = 15 - bb;
        outv[j] = dense ? accr[c] : accg[cp + j];
        outc[j] = (unsigned char)c;
    }
}

// ------------------------------------------------------------------ step 3, element-wise
// Segments of tile-rows the row-wise tile kernel cannot hold (too many C
// tiles, A tiles or B tiles, typically hypersparse rows where most tiles
// carry one or two nonzeros): element-level Gustavson with B's row index.
// One 256-thread block per segment (a run of C tiles holding <= G3E_NNZ
// nonzeros, accumulated in LDS in final order and located through an LDS
// hash of the segment's C entries); 16 groups of 16 lanes, group r owns C
// row 16I+r.  The group walks A row 16I+r in ascending k (A tiles in
// ascending K, bits in ascending column) and, for each a_ik, its lanes
// scatter the entries of B row k that fall in the segment's column range:
// distinct C entries per step, steps in ascending k, so each C entry again
// sums its products in the serial SPA order, without atomics.  Segments of
// 1536 nonzeros balance the re-walks of A rows per tile-row against the
// block's LDS (accumulator, hash, column bytes), which sets the blocks per
// CU in gfx950's 1,280-byte LDS granules.
#define G3E_NNZ 1536

// LDS hash of the segment's C entries: key (column - col_lo) * 16 + row,
// stored + 1 (0 = empty), value = the entry's position in the segment,
// linear probing.  2560 slots (load factor at most 0.6): six blocks per CU
// packed (25.6 KB of LDS; 24.2 KB with A's row index, which needs no
// per-group decode scratch), five unpacked when A is decoded from its tiles
// (30.7 KB).  Unpacked with A's row index, 2048 slots (at most 0.75) fit
// six blocks (26.1 KB); the packed table stays at 2560, since at seven
// blocks per CU its higher load factor cost more than the occupancy gave.
#define G3E_EHASH 2560
#define G3E_EHASH_RXW 2048
#define G3E_PF 4  // C tiles per thread whose metadata is loaded per round
// home slot of a key in a table of HS slots (multiply-shift range
// reduction) and the next slot
template <unsigned HS>
__device__ __forceinline__ unsigned ent_slot(unsigned key) {
    return (unsigned)(((unsigned long long)(key * 2654435761u) * HS) >> 32);
}
template <unsigned HS>
__device__ __forceinline__ unsigned ent_next(unsigned h) { return h + 1 == HS ? 0u : h + 1; }
// PK: when every tile-row spans fewer than 8192 tile columns the key (+1) is
// below 2^21 and one 32-bit slot holds key << 11 | position (positions are
// below G3E_NNZ <= 2048); without the separate position array the block's
// LDS fits six blocks per CU.
#define G3E_PK_SPAN 8192
static_assert(G3E_NNZ <= 2048, "positions take 11 bits of a packed hash slot");

// RX: A carries its element row index (rows' entries read directly); else
// A row r's entries are decoded from its tiles.
template <bool PK, bool RX>
__global__ void __launch_bounds__(256, PK || RX ? 6 : 5)
k_step3_elem(TileOperands op, const int *__restrict__ Browp, const int *__restrict__ Bcolx,
             const double *__restrict__ Bvalx, int m, const int4 *__restrict__ segs, const int *__restrict__ nseg,
             const int *__restrict__ Cptr, const int *__restrict__ Ccol, const unsigned short *__restrict__ Cmask,
             const unsigned char *__restrict__ Cptr16, const int *__restrict__ Cnnz, unsigned char *__restrict__ Ccol8,
             double *__restrict__ Cval, const unsigned char *__restrict__ heavy) {
    __shared__ double s_acc[G3E_NNZ];
    constexpr unsigned HS = RX && !PK ? G3E_EHASH_RXW : G3E_EHASH;
    __shared__ unsigned s_ekey[HS];
    __shared__ unsigned short s_epos[PK ? 1 : HS];
    __shared__ unsigned char s_ecol[G3E_NNZ];  // in-tile column of each entry, final order
    const int tid = threadIdx.x;
    // the grid is an upper bound: the first *nseg blocks (spread evenly over
    // the XCDs) take the segments, each XCD a contiguous range
    const int ns = *nseg;
    if ((int)blockIdx.x >= ns) return;  // block-uniform
    const int w = (int)tsg_xcd_block(blockIdx.x, (unsigned)ns);
    const int4 sg = segs[w];
    const int I = sg.x;
    const int c0 = Cptr[I] + sg.y, u = sg.z - sg.y;
    const int nbase = Cnnz[c0], nseg_nnz = Cnnz[c0 + u] - nbase;
    const int col_lo = Ccol[c0] * 16, col_hi = Ccol[c0 + u - 1] * 16 + 16;
    for (int i = tid; i < nseg_nnz; i += 256) s_acc[i] = 0.0;
    for (int i = tid; i < (int)HS; i += 256) s_ekey[i] = 0u;
    __syncthreads();
    // every C entry of the segment into the hash, and its in-tile column
    // into s_ecol (one C tile per thread; structurally empty tiles skipped).
    // The metadata of up to G3E_PF of a thread's tiles is loaded in one
    // round before their entries are inserted.
    for (int i0 = 0; i0 < u; i0 += 256 * G3E_PF) {
        int pos0[G3E_PF], nt[G3E_PF];
        unsigned j16[G3E_PF];
        uint4 mA[G3E_PF], mB[G3E_PF];
#pragma unroll
        for (int v = 0; v < G3E_PF; v++) {
            const int i = i0 + v * 256 + tid;
            nt[v] = 0;
            if (i < u) {
                const long long t = c0 + i;
                pos0[v] = Cnnz[t] - nbase;
                nt[v] = Cnnz[t + 1] - Cnnz[t];
                const uint4 *cm4 = reinterpret_cast<const uint4 *>(Cmask + t * 16);
                mA[v] = cm4[0];
                mB[v] = cm4[1];
                j16[v] = (unsigned)(Ccol[t] * 16 - col_lo);
            }
        }
#pragma unroll
        for (int v = 0; v < G3E_PF; v++) {
            if (nt[v] == 0) continue;
            const unsigned mw[8] = {mA[v].x, mA[v].y, mA[v].z, mA[v].w, mB[v].x, mB[v].y, mB[v].z, mB[v].w};
            int pos = pos0[v];
            for (int rr = 0; rr < 16; rr++) {  // rows in order: positions follow
                unsigned cm = (mw[rr >> 1] >> ((rr & 1) * 16)) & 0xffffu;
                while (cm) {
                    const int bb = 31 - __clz(cm);
                    cm &= ~(1u << bb);
                    const unsigned key = (((j16[v] + 15 - bb) << 4) | (unsigned)rr) + 1u;
                    unsigned h = ent_slot<HS>(key);
                    if (PK) {
                        const unsigned ent = key << 11 | (unsigned)pos;
                        while (atomicCAS(&s_ekey[h], 0u, ent) != 0u) h = ent_next<HS>(h);
                    } else {
                        while (atomicCAS(&s_ekey[h], 0u, key) != 0u) h = ent_next<HS>(h);  // nnz < HS
                        s_epos[h] = (unsigned short)pos;
                    }
                    s_ecol[pos] = (unsigned char)(15 - bb);
                    pos++;
                }
            }
        }
    }
    __syncthreads();
    const int r = tid >> 4, l = tid & 15;
    const int row = 16 * I + r;
    // rows flagged heavy are computed by k_step3_heavy, which runs after
    // this kernel and overwrites their values
    const bool skip = heavy && row < m && heavy[row];
    const int gb = tid & 48;  // first lane of this group in its wave
    __shared__ int s_gex[RX ? 1 : 16][16];             // per group: entries before each lane's tile
    __shared__ unsigned short s_gam[RX ? 1 : 16][16];  // per group: each lane's tile row mask
    // adds a * b into C entry (row 16I+r, col), located through the hash
    // (the entry is in C's structure; the probe bound is not reached for a
    // consistent C)
    auto acc_add = [&](int col, double a, double b) {
        const unsigned key = ((unsigned)(col - col_lo) << 4 | (unsigned)r) + 1u;
        unsigned h = ent_slot<HS>(key);
        int probe = 0;
        if (PK) {
            unsigned ent;
            while (((ent = s_ekey[h]) >> 11) != key && ++probe < (int)HS) h = ent_next<HS>(h);
            if (probe < (int)HS) {
                double *acc = &s_acc[ent & 2047u];
                *acc = __builtin_fma(a, b, *acc);
            }
        } else {
            while (s_ekey[h] != key && ++probe < (int)HS) h = ent_next<HS>(h);
            if (probe < (int)HS) {
                double *acc = &s_acc[s_epos[h]];
                *acc = __builtin_fma(a, b, *acc);
            }
        }
    };
    if (RX && row < m && nseg_nnz > 0 && !skip) {
        // A's row index: row 16I+r's entries are contiguous, ascending k.  The
        // 16 lanes take 16 of them (sources) and find their B ranges in
        // parallel; then the first 16 in-range B entries of all 16 sources are
        // loaded at once (lane l holds entry l of each), so the ordered
        // accumulation below waits on no global load for rows of up to 16
        // entries in the segment's column range.
        const int x1 = op.Arowp[row + 1];
        for (int x0 = op.Arowp[row]; x0 < x1; x0 += 16) {
            const int x = x0 + l;
            int e0 = 0, e1 = 0;
            double a = 0.0;
            if (x < x1) {
                const int k = op.Acolx[x];
                a = op.Avalx[x];
                e1 = Browp[k + 1];
                // every entry of B row k is in a C tile of the tile-row, so
                // none lies before the row's first segment
                e0 = sg.w == 0 ? Browp[k] : lower_bound_i(Bcolx, Browp[k], e1, col_lo);
            }
            int pc[16];
            double pv[16];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int qs0 = tsg_row16_bcast(e0, q), qs1 = tsg_row16_bcast(e1, q);
                const int e = qs0 + l;
                pc[q] = 0x7fffffff;
                pv[q] = 0.0;
                if (e < qs1) {
                    pc[q] = Bcolx[e];
                    pv[q] = Bvalx[e];
                }
            }
#pragma unroll
            for (int q = 0; q < 16; q++) {  // sources in ascending k
                const double as = tsg_row16_bcast(a, q);
                const int qs0 = tsg_row16_bcast(e0, q), qs1 = tsg_row16_bcast(e1, q);
                if (pc[q] < col_hi) acc_add(pc[q], as, pv[q]);
                // entries past the first 16 (the last prefetched one in range)
                if (qs1 - qs0 > 16 && tsg_row16_bcast(pc[q], 15) < col_hi) {
                    for (int e = qs0 + 16 + l; e < qs1; e += 16) {
                        const int col = Bcolx[e];
                        if (col >= col_hi) break;
                        acc_add(col, as, Bvalx[e]);
                    }
                }
                tsg_wave_sync();  // the next k of this row reads what lanes wrote
            }
        }
    } else if (!RX && row < m && nseg_nnz > 0 && !skip) {
        const int a0 = op.Aptr[I], a1 = op.Aptr[I + 1];
        for (int t0 = a0; t0 < a1; t0 += 16) {
            // 16 A tiles per round, one per lane: their row-r entries are
            // numbered x = 0..total-1 in ascending k (tiles in ascending K,
            // columns ascending inside a tile)
            const int tl = t0 + l;
            const unsigned am_l = tl < a1 ? op.Amask[(long long)tl * 16 + r] : 0u;
            const int cnt = __popc(am_l);
            int incl = cnt;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const int v = __shfl_up(incl, o, 16);
                if (l >= o) incl += v;
            }
            const int total = __shfl(incl, 15, 16);
            s_gex[r][l] = incl - cnt;
            s_gam[r][l] = (unsigned short)am_l;
            tsg_wave_sync();
            for (int x0 = 0; x0 < total; x0 += 16) {
                // lane l takes entry x0 + l and finds its B range in parallel
                const int x = x0 + l;
                int k = 0, e0 = 0, e1 = 0;
                double a = 0.0;
                bool have = false;
                if (x < total) {
                    int ow = 0;
#pragma unroll
                    for (int st = 8; st > 0; st >>= 1)
                        if (ow + st < 16 && s_gex[r][ow + st] <= x) ow += st;
                    unsigned am = s_gam[r][ow];
                    int j = x - s_gex[r][ow];  // j-th entry of the owner's row
                    for (int q = 0; q < j; q++) am &= ~(1u << (31 - __clz(am)));
                    const int b = 31 - __clz(am);
                    const int t = t0 + ow;
                    k = op.Acol[t] * 16 + 15 - b;
                    a = op.Aval[op.Annz[t] + op.Aptr16[(long long)t * 16 + r] + j];
                    e1 = Browp[k + 1];
                    e0 = lower_bound_i(Bcolx, Browp[k], e1, col_lo);
                    have = e0 < e1 && Bcolx[e0] < col_hi;
                }
                unsigned hv = (unsigned)(__ballot(have) >> gb) & 0xffffu;
                while (hv) {  // entries with B entries in range, ascending k
                    const int src = __ffs(hv) - 1;
                    hv &= hv - 1;
                    const double as = __shfl(a, gb + src, 64);
                    const int es0 = __shfl(e0, gb + src, 64), es1 = __shfl(e1, gb + src, 64);
                    for (int e = es0 + l; e < es1; e += 16) {
                        const int col = Bcolx[e];
                        if (col >= col_hi) break;
                        acc_add(col, as, Bvalx[e]);
                    }
                    tsg_wave_sync();  // the next k of this row reads what lanes wrote
                }
            }
            tsg_wave_sync();  // s_gex / s_gam are rewritten next round
        }
    }
    __syncthreads();
    for (int i = tid; i < nseg_nnz; i += 256) {
        Cval[(long long)nbase + i] = s_acc[i];
        Ccol8[(long long)nbase + i] = s_ecol[i];
    }
}

// ------------------------------------------------------------------ step 3, heavy rows
// Rows with very many intermediate products (hub rows of power-law graphs)
// would make every C segment of their tile-row re-walk the whole A row.
// They are computed one wave at a time instead, in column parts of
// HEAVY_PART_COLS columns whose FP64 accumulators live in the wave's own
// LDS window: for part [lo, lo + HEAVY_PART_COLS) the wave walks A row i in
// ascending k, its lanes adding a_ik * b_kj for the entries of B row k in
// the part (distinct columns per k; LDS instructions of one wave complete
// in order, so the next k reads what this one wrote), then copies the part
// into C's final positions.  A work item is a run of consecutive parts of
// one row; between parts the wave keeps, per A entry, the position in
// B row k where the next part starts (a cursor in global scratch), so only
// an item's first part searches B's rows.
#define HEAVY_MIN_PRODUCTS 65536
#ifndef HEAVY_PART_COLS
#define HEAVY_PART_COLS 2048         // 16 KiB of LDS per wave, seven waves per CU
#endif
#define HEAVY_ITEM_PRODUCTS 32768    // products per work item, about
#define HEAVY_MAX_WAVES 4096
#define HEAVY_CURSOR_BYTES (512ll << 20)

// Per C row of an element-wise tile-row: heavy flag and work items (runs of
// parts); info[0] = longest heavy A row, info[1] = heavy rows.
__global__ void __launch_bounds__(256)
k_heavy_flags(const int *__restrict__ Arowp, const int *__restrict__ Acolx, const int *__restrict__ Browp,
              const unsigned char *__restrict__ rowcls, const int *__restrict__ Cptr, const int *__restrict__ Ccol,
              int m, long long min_products, long long item_products, unsigned char *__restrict__ flag,
              int *__restrict__ cnt, int *__restrict__ info) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    int h = 0, items = 0;
    if (i < m && rowcls[i >> 4] == 0) {
        const int x0 = Arowp[i], x1 = Arowp[i + 1];
        long long p = -1;  // unknown for long A rows
        if (x1 - x0 > 2048) {
            h = 1;
        } else {
            p = 0;
            for (int x = x0; x < x1; x++) p += Browp[Acolx[x] + 1] - Browp[Acolx[x]];
            h = p > min_products;
        }
        if (h) {
            const int I = i >> 4, c0 = Cptr[I], c1 = Cptr[I + 1];
            const long long span = (long long)(Ccol[c1 - 1] - Ccol[c0] + 1) * 16;
            const int nparts = (int)((span + HEAVY_PART_COLS - 1) / HEAVY_PART_COLS);
            items = p < 0 ? nparts
                          : (int)min<long long>(nparts, max<long long>(1, (p + item_products - 1) / item_products));
            atomicMax(&info[0], x1 - x0);
            atomicAdd(&info[1], 1);
        }
    }
    if (i < m) {
        flag[i] = (unsigned char)h;
        cnt[i] = items;
    }
}

// Work items {row, first part | end part << 16}, parts split evenly.
__global__ void __launch_bounds__(256)
k_heavy_list(const int *__restrict__ off, const unsigned char *__restrict__ flag, const int *__restrict__ Cptr,
             const int *__restrict__ Ccol, int m, int2 *__restrict__ items) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m || !flag[i]) return;
    const int I = i >> 4, c0 = Cptr[I], c1 = Cptr[I + 1];
    const long long span = (long long)(Ccol[c1 - 1] - Ccol[c0] + 1) * 16;
    const int nparts = (int)((span + HEAVY_PART_COLS - 1) / HEAVY_PART_COLS);
    const int o = off[i], n = off[i + 1] - o;
    for (int g = 0; g < n; g++) {
        const int p0 = (int)((long long)nparts * g / n), p1 = (int)((long long)nparts * (g + 1) / n);
        items[o + g] = make_int2(i, (int)((unsigned)p0 | ((unsigned)p1 << 16)));
    }
}

#ifndef HEAVY_STAGE
#define HEAVY_STAGE 8  // B entries per source staged in LDS per part
#endif
// lane q's double, read by the whole wave (q wave-uniform)
__device__ __forceinline__ double rl_f64(double v, int q) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, q);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), q);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__global__ void __launch_bounds__(64)
k_step3_heavy(const int *__restrict__ Arowp, const int *__restrict__ Acolx, const double *__restrict__ Avalx,
              const int *__restrict__ Browp, const int *__restrict__ Bcolx, const double *__restrict__ Bvalx,
              const int2 *__restrict__ items, const int *__restrict__ nitems, const int *__restrict__ Cptr,
              const int *__restrict__ Ccol, const unsigned short *__restrict__ Cmask,
              const unsigned char *__restrict__ Cptr16, const int *__restrict__ Cnnz, double *__restrict__ Cval,
              int *__restrict__ cursor, long long cur_stride, int *__restrict__ next_item) {
    __shared__ double s_acc[HEAVY_PART_COLS];
    __shared__ double s_sv[64 * HEAVY_STAGE];  // staged B values, source-major
    __shared__ int s_sc[64 * HEAVY_STAGE];     // their columns
    const int lane = threadIdx.x;
    int *cur = cursor + (long long)blockIdx.x * cur_stride;
    const int ni = *nitems;
    // items are taken from a shared counter, so waves that drew short items
    // take more (every wave exits once the counter passes ni)
    for (;;) {
        int h = 0;
        if (lane == 0) h = atomicAdd(next_item, 1);
        h = __shfl(h, 0, 64);
        if (h >= ni) break;
        const int2 it = items[h];
        const int i = it.x, p0 = (int)((unsigned)it.y & 0xffffu), p1 = (int)((unsigned)it.y >> 16);
        const int I = i >> 4, r = i & 15;
        const int c0 = Cptr[I], c1 = Cptr[I + 1];
        const int colbase = Ccol[c0] * 16;
        const int xa = Arowp[i], xe = Arowp[i + 1];
        for (int part = p0; part < p1; part++) {
            const int lo = colbase + part * HEAVY_PART_COLS, hi = lo + HEAVY_PART_COLS;
            for (int j = lane; j < HEAVY_PART_COLS; j += 64) s_acc[j] = 0.0;
            for (int xc = xa; xc < xe; xc += 64) {
                // 64 A entries (sources) at once: lane q finds where B row k
                // enters the part and stages its first HEAVY_STAGE entries
                const int n = min(64, xe - xc);
                int cb = 0, ce = 0, cnt = 0;
                bool more = false;
                double aq = 0.0;  // lane q: a_ik of source q
                if (lane < n) {
                    const int k = Acolx[xc + lane];
                    aq = Avalx[xc + lane];
                    ce = Browp[k + 1];
                    cb = part == p0 ? lower_bound_i(Bcolx, Browp[k], ce, lo) : cur[xc - xa + lane];
                    int col[HEAVY_STAGE];
                    double val[HEAVY_STAGE];
#pragma unroll
                    for (int j = 0; j < HEAVY_STAGE; j++) {
                        col[j] = cb + j < ce ? Bcolx[cb + j] : 0x7fffffff;
                        val[j] = cb + j < ce ? Bvalx[cb + j] : 0.0;
                    }
#pragma unroll
                    for (int j = 0; j < HEAVY_STAGE; j++) {
                        const bool in = col[j] < hi;  // in-range entries form a prefix
                        cnt += in;
                        if (in) {
                            s_sc[lane * HEAVY_STAGE + j] = col[j];
                            s_sv[lane * HEAVY_STAGE + j] = val[j];
                        }
                    }
                    more = cnt == HEAVY_STAGE && cb + HEAVY_STAGE < ce;
                }
                // lane q: staged entries of source q | more may follow << 8;
                // read by the whole wave with v_readlane, not through LDS
                const int nq = cnt | (more ? 256 : 0);
                int next = cb + cnt;  // lane q: where source q's next part starts
                // only sources with entries in the part take a step
                unsigned long long act = __ballot(lane < n && cnt > 0);
                tsg_wave_sync();
                // the staged data of the next active source is read before this
                // source's accumulation, which then costs one LDS round trip;
                // so are the first 256 entries past a source's staged ones
                // (global loads into registers, one source ahead)
                auto tail_first = [&](int qq, int snq, int *tc, double *tv) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        tc[u] = 0x7fffffff;
                        tv[u] = 0.0;
                    }
                    if (snq & 256) {
                        const int e = __builtin_amdgcn_readlane(cb, qq) + HEAVY_STAGE;
                        const int e1 = __builtin_amdgcn_readlane(ce, qq);
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int ee = e + 64 * u + lane;
                            if (ee < e1) {
                                tc[u] = Bcolx[ee];
                                tv[u] = Bvalx[ee];
                            }
                        }
                    }
                };
                int q = act ? __ffsll((long long)act) - 1 : 0;
                int sn = act ? __builtin_amdgcn_readlane(nq, q) : 0, sc = 0;
                double a = act ? rl_f64(aq, q) : 0.0, sv = 0.0;
                if (lane < (sn & 255)) {
                    sc = s_sc[q * HEAVY_STAGE + lane];
                    sv = s_sv[q * HEAVY_STAGE + lane];
                }
                int tc[4];
                double tv[4];
                tail_first(q, sn, tc, tv);
                while (act) {  // ascending k; entries of one k have distinct columns
                    act &= act - 1;
                    const int q1 = act ? __ffsll((long long)act) - 1 : 0;
                    int sn1 = 0, sc1 = 0;
                    double a1 = 0.0, sv1 = 0.0;
                    if (act) {
                        sn1 = __builtin_amdgcn_readlane(nq, q1);
                        a1 = rl_f64(aq, q1);
                        if (lane < (sn1 & 255)) {
                            sc1 = s_sc[q1 * HEAVY_STAGE + lane];
                            sv1 = s_sv[q1 * HEAVY_STAGE + lane];
                        }
                    }
                    int tc1[4];
                    double tv1[4];
                    tail_first(q1, sn1, tc1, tv1);
                    if (lane < (sn & 255)) s_acc[sc - lo] = __builtin_fma(a, sv, s_acc[sc - lo]);
                    if (sn & 256) {  // entries past the staged ones, 256 per round trip
                        int e = __builtin_amdgcn_readlane(cb, q) + HEAVY_STAGE;
                        const int e1 = __builtin_amdgcn_readlane(ce, q);
                        bool go;
                        {  // the first round, loaded one source ahead
                            int got = 0;
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const bool in2 = tc[u] < hi;
                                if (in2) s_acc[tc[u] - lo] = __builtin_fma(a, tv[u], s_acc[tc[u] - lo]);
                                got += __popcll(__ballot(in2));
                            }
                            e += got;
                            go = got == 256;  // the part may hold more of this row
                        }
                        while (go && e < e1) {
                            int c[4];
                            double v[4];
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const int ee = e + 64 * u + lane;
                                c[u] = ee < e1 ? Bcolx[ee] : 0x7fffffff;
                                v[u] = ee < e1 ? Bvalx[ee] : 0.0;
                            }
                            int got = 0;
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const bool in2 = c[u] < hi;
                                if (in2) s_acc[c[u] - lo] = __builtin_fma(a, v[u], s_acc[c[u] - lo]);
                                got += __popcll(__ballot(in2));
                            }
                            e += got;
                            if (got < 256) break;
                        }
                        if (lane == q) next = e;
                    }
                    tsg_wave_sync();  // the next k reads what this one wrote
                    q = q1;
                    sn = sn1;
                    sc = sc1;
                    a = a1;
                    sv = sv1;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        tc[u] = tc1[u];
                        tv[u] = tv1[u];
                    }
                }
                if (part + 1 < p1 && lane < n) cur[xc - xa + lane] = next;
            }
            // the part's C entries of row i, into their final positions
            const int t0 = lower_bound_i(Ccol, c0, c1, lo >> 4), t1 = lower_bound_i(Ccol, t0, c1, hi >> 4);
            for (int t = t0 + lane; t < t1; t += 64) {
                unsigned cm = Cmask[(long long)t * 16 + r];
                long long pos = (long long)Cnnz[t] + Cptr16[(long long)t * 16 + r];
                const double *d = s_acc + (Ccol[t] * 16 - lo);
                while (cm) {
                    const int bb = 31 - __clz(cm);
                    cm &= ~(1u << bb);
                    Cval[pos++] = d[15 - bb];
                }
            }
            tsg_wave_sync();  // the next part clears the window
        }
    }
}

// ------------------------------------------------------------------ step 3, row-wise
// One 256-thread block per C tile-row whose C tiles, C nonzeros, A tiles and
// B tile-rows fit in LDS (g3 class, decided by k_row_class).  The
// tile-row's C values accumulate in LDS in their final order (tile by tile,
// row by row, ascending column), so the write-out is one contiguous copy.
// A tiles (I,K) are taken in ascending K; every (B tile q, nonempty A row
// r) task multiplies A row r by the B tile into C tile slot(q), row r.
// C row r belongs to wave g3_owner(r): only that wave ever adds into it,
// and each wave takes its tasks in ascending K, so every C entry receives its
// products in ascending k -- the order of the serial SPA -- with no atomics
// and no barrier between the K-steps for the accumulator's sake.
//
// An A tile with more than one entry has its A tile and its whole B
// tile-row K -- contiguous in row-major tile order -- staged into LDS with
// coalesced loads; an A tile with a single entry (r, k) needs row k of each
// B tile only, which its row's wave loads into registers itself (the
// stencil corner tiles: 18 of the 27 A tiles of a 27-point tile-row).
//
// The staging is software-pipelined: the tile-row's (K, B tile-row) table
// is loaded once, a staged K-step sits in the second LDS buffer while the
// previous one is computed, and the global loads of the step after next
// are in flight meanwhile.  A barrier precedes each staged K-step only.
#define G3_LAMAX 64  // A tiles per tile-row (row table built by wave 0)
// A C tile with at least G3_PADMIN nonzeros is followed by one unused
// accumulator element, so that the starts of equally sized tiles (a
// stencil's repeating 3 / 74 / 3 nonzeros) fall in different LDS banks; the
// lanes of a wave update the same row of many such tiles at once.
#define G3_PADMIN 16

// Dynamic LDS layout of k_step3_rows, from the maxima over its rows.
struct G3Layout {
    int umax, nnzmax, lamax, qb, bnnz, amax;
};
#define G3_NARR 13
__host__ __device__ inline size_t g3_layout(const G3Layout &L, size_t *off) {
    const size_t sz[G3_NARR] = {
        8u * (size_t)L.nnzmax,            // 0 acc
        // 1 B values (2 stages); before the first stage, the C tile columns
        // (int) and their slot ids (u8) of all the row's C tiles
        2u * 8u * (size_t)L.bnnz > 5u * G3_UMAX ? 2u * 8u * (size_t)L.bnnz : 5u * G3_UMAX,
        2u * 8u * (size_t)L.amax,         // 2 A values
        (size_t)L.umax,                   // 3 accumulator pads before each C tile (u8)
        2u * (size_t)L.umax,              // 4 C tile offsets (u16)
        32u * (size_t)L.umax,             // 5 C masks
        16u * (size_t)L.umax,             // 6 C row offsets
        20u * (size_t)L.lamax,            // 7 row table (5 ints per A tile)
        (size_t)L.lamax * (size_t)L.qb,   // 8 pair slots
        2u * 32u * (size_t)L.qb,          // 9 B masks
        2u * 16u * (size_t)L.qb,          // 10 B row offsets
        2u * 2u * ((size_t)L.qb + 1),     // 11 B tile nnz offsets (u16)
        2u * 2u * (size_t)L.qb};          // 12 B nonempty-row bits
    size_t o = 0;
    for (int i = 0; i < G3_NARR; i++) {
        off[i] = o;
        o += (sz[i] + 15) & ~(size_t)15;
    }
    return o;
}

// Pipeline stage registers.  Staging is split by wave so every branch is
// wave-uniform: wave 0 loads the B tile masks (uint4 = 8 rows, 2 per tile),
// row offsets (uint4 per tile) and nnz offsets; wave 1 the A tile (values,
// masks, row offsets); waves 2-3 the B values (4 per lane).  The ring has
// two register slots and runs over the staged (multi-entry) K-steps only: a
// stage is loaded two staged steps before it is computed and stored into
// LDS one staged step before.  The roles share one set of 10 registers per
// slot: wave 0 keeps its two 16-byte words as raw bits in d0:d1 and d2:d3
// (moved, never computed on).
#define G3_SLOT_DECL(S)                                                   \
    double S##d0 = 0.0, S##d1 = 0.0, S##d2 = 0.0, S##d3 = 0.0;            \
    unsigned S##e0 = 0u, S##e1 = 0u

#define G3_LOAD(j, S)                                                                                         \
    do {                                                                                                      \
        const int lb_ = G3_LB(j), q0_ = SROW(3, j);                                                           \
        /* Every lane loads from an in-bounds address (clamped; the lanes past  \
           the stage's data load a valid element and the store skips them), so \
           the loads are straight-line code and the store of this stage, two   \
           K-steps later, waits for these loads only. */                       \
        if (wave == 0) {                                                                                      \
            const long long bq_ = lb_ > 0 ? (long long)q0_ * 16 : 0;                                          \
            const double2 m_ = reinterpret_cast<const double2 *>(op.Bmask + bq_)[min(lane, max(2 * lb_ - 1, 0))]; \
            const double2 o_ = reinterpret_cast<const double2 *>(op.Bptr16 + bq_)[min(lane, max(lb_ - 1, 0))]; \
            S##d0 = m_.x; S##d1 = m_.y; S##d2 = o_.x; S##d3 = o_.y;                                           \
            S##e0 = op.Bnnz[q0_ + min(lane, lb_)]; /* made relative at the store */                           \
        } else if (wave == 1) {                                                                               \
            const int pan_ = G3_PAN(j), pa0_ = SROW(1, j); /* A tiles are never empty */                    \
            S##d0 = op.Aval[pa0_ + min(lane, pan_ - 1)];                                                      \
            S##d1 = op.Aval[pa0_ + min(lane + 64, pan_ - 1)];                                                 \
            S##d2 = op.Aval[pa0_ + min(lane + 128, pan_ - 1)];                                                \
            S##d3 = op.Aval[pa0_ + min(lane + 192, pan_ - 1)];                                                \
            const long long x_ = (long long)(a0 + (j)) * 16 + (lane & 15); /* A tile a0 + j */              \
            S##e0 = op.Amask[x_];                                                                             \
            S##e1 = op.Aptr16[x_];                                                                            \
        } else {                                                                                              \
            const int pbn_ = SROW(5, j), b_ = (wave - 2) * 64 + lane;                                         \
            const int pb0_ = pbn_ > 0 ? SROW(4, j) : 0, pbl_ = max(pbn_ - 1, 0);                              \
            S##d0 = op.Bval[pb0_ + min(b_, pbl_)];                                                            \
            S##d1 = op.Bval[pb0_ + min(b_ + 128, pbl_)];                                                      \
            S##d2 = op.Bval[pb0_ + min(b_ + 256, pbl_)];                                                      \
            S##d3 = op.Bval[pb0_ + min(b_ + 384, pbl_)];                                                      \
        }                                                                                                     \
    } while (0)

#define G3_STORE(j, S, buf)                                                                                   \
    do {                                                                                                      \
        const int lb_ = G3_LB(j);                                                                             \
        if (wave == 0) {                                                                                      \
            if (lane < 2 * lb_) reinterpret_cast<double2 *>(s_bm + (buf) * qb * 16)[lane] = make_double2(S##d0, S##d1); \
            if (lane < lb_) reinterpret_cast<double2 *>(s_bp + (buf) * qb * 16)[lane] = make_double2(S##d2, S##d3); \
            if (lane <= lb_) s_bnnz[(buf) * (qb + 1) + lane] = (unsigned short)((int)S##e0 - SROW(4, j));     \
            /* nonempty-row bits of this half tile, OR-ed with the partner half */                            \
            const unsigned wd_[4] = {(unsigned)__double2loint(S##d0), (unsigned)__double2hiint(S##d0),        \
                                     (unsigned)__double2loint(S##d1), (unsigned)__double2hiint(S##d1)};       \
            unsigned nz_ = 0;                                                                                 \
            _Pragma("unroll") for (int i_ = 0; i_ < 4; i_++) {                                                \
                const int k_ = (lane & 1) * 8 + 2 * i_;                                                       \
                nz_ |= ((wd_[i_] & 0xffffu) ? 1u : 0u) << (15 - k_);                                          \
                nz_ |= ((wd_[i_] >> 16) ? 1u : 0u) << (14 - k_);                                              \
            }                                                                                                 \
            nz_ |= __shfl_xor(nz_, 1, 64);                                                                    \
            if (!(lane & 1) && lane < 2 * lb_) s_bnz[(buf) * qb + (lane >> 1)] = (unsigned short)nz_;         \
        } else if (wave == 1) {                                                                               \
            const int pan_ = G3_PAN(j);                                                                      \
            double *av_ = s_aval + (buf) * amax;                                                              \
            if (lane < pan_) av_[lane] = S##d0;                                                               \
            if (lane + 64 < pan_) av_[lane + 64] = S##d1;                                                     \
            if (lane + 128 < pan_) av_[lane + 128] = S##d2;                                                   \
            if (lane + 192 < pan_) av_[lane + 192] = S##d3;                                                   \
            if (lane < 16) {                                                                                  \
                s_am[buf][lane] = (unsigned short)S##e0;                                                      \
                s_ap[buf][lane] = (unsigned char)S##e1;                                                       \
            }                                                                                                 \
        } else {                                                                                              \
            const int pbn_ = SROW(5, j), b_ = (wave - 2) * 64 + lane;                                         \
            double *bv_ = s_bval + (buf) * bnnzs;                                                             \
            if (b_ < pbn_) bv_[b_] = S##d0;                                                                   \
            if (b_ + 128 < pbn_) bv_[b_ + 128] = S##d1;                                                       \
            if (b_ + 256 < pbn_) bv_[b_ + 256] = S##d2;                                                       \
            if (b_ + 384 < pbn_) bv_[b_ + 384] = S##d3;                                                       \
        }                                                                                                     \
    } while (0)

// Row table entry k of A tile j (k = 1..5): 1 its first value, 2 its entry
// count | the tile count of B tile-row K << 9 | the rank of its first (A
// tile, B tile) pair << 15 (at most 256, 32 and 63 * 32), 3 the first tile
// of B tile-row K, 4 and 5 the first value and value count of that B
// tile-row (for a single-entry tile: 4 = its row | column << 8).
#define SROW(k, j) s_row[((k) - 1) * lamax + (j)]
#define G3_PAN(j) (SROW(2, j) & 511)
#define G3_LB(j) ((SROW(2, j) >> 9) & 63)
#define G3_PB(j) (SROW(2, j) >> 15)

// Wave owning C row r: rows r = 4i + c belong to wave (2, 0, 1, 3)[c], so the
// rows whose single-entry tiles a stencil produces (0 and 15) fall on the
// waves that stage only B values; g3_rbase(w) = the c of wave w.
__device__ __forceinline__ int g3_owner(int r) { return (0xD2 >> (2 * (r & 3))) & 3; }
__device__ __forceinline__ int g3_rbase(int w) { return (0xC9 >> (2 * w)) & 3; }

// Single-entry K-steps of one wave, lane = B tile, in a two-level pipeline
// over the steps the wave owns: the step after next has row k's mask, its
// tile's value offset and row k's offset within the tile in flight (G3_L1:
// sm_*); the next has row k's first two values, the A value, and its C
// row's mask and accumulator offset from the pair's slot (G3_L2: sv_*).
// Every address is in bounds (a lane past the B tile-row or with an empty
// row k loads a valid element and computes nothing).
#define G3_L1(j)                                                                      \
    do {                                                                              \
        const int rk_ = SROW(4, j), nq_ = G3_LB(j);                                  \
        const int q_ = max(SROW(3, j) + min(lane, nq_ - 1), 0);                       \
        const long long x_ = (long long)q_ * 16 + (rk_ >> 8);                         \
        sm_m = op.Bmask[x_];                                                          \
        sm_n = (unsigned)op.Bnnz[q_];                                                 \
        sm_p = op.Bptr16[x_];                                                         \
    } while (0)
#define G3_L2(j)                                                                      \
    do {                                                                              \
        const int nb_ = __popc(sm_m), nq_ = G3_LB(j), rr_ = SROW(4, j) & 15;         \
        const int pos_ = (int)(sm_n + sm_p), bs_ = nb_ > 0 ? pos_ : (int)sm_n;        \
        sv_0 = op.Bval[bs_];                                                          \
        sv_1 = op.Bval[bs_ + min(1, max(nb_ - 1, 0))];                                \
        sv_a = op.Aval[SROW(1, j)];                                                   \
        sv_m = sm_m;                                                                  \
        sv_p = pos_;                                                                  \
        const int sl_ = nq_ > 0 ? s_pslot[G3_PB(j) + min(lane, nq_ - 1)] : 0;         \
        sv_cm = s_cmask[sl_ * 16 + rr_];                                              \
        sv_cb = s_coff[sl_] + s_pad[sl_] + s_cptr[sl_ * 16 + rr_];                    \
    } while (0)
// K-step j = (row rr, column k): row k of the lane's B tile into row rr of
// the pair's C tile, ascending column (values past the second from memory);
// then the pipeline moves one step on
#define G3_SINGLE(j)                                                                  \
    do {                                                                              \
        unsigned bm_ = sv_m;                                                          \
        if (lane < G3_LB(j) && bm_) {                                                 \
            double *ac_ = s_acc + sv_cb;                                              \
            int jb_ = 0;                                                              \
            while (bm_) {                                                             \
                const int bb_ = 31 - __clz(bm_);                                      \
                bm_ &= ~(1u << bb_);                                                  \
                const double b_ = jb_ == 0 ? sv_0 : jb_ == 1 ? sv_1 : op.Bval[sv_p + jb_]; \
                double *acc_ = ac_ + __popc(sv_cm >> (bb_ + 1));                      \
                *acc_ = __builtin_fma(sv_a, b_, *acc_);                               \
                jb_++;                                                                \
            }                                                                         \
        }                                                                             \
        sA = sB;                                                                      \
        if (sA < 64) G3_L2(sA);                                                       \
        sB = g3_pop(own);                                                             \
        if (sB < 64) G3_L1(sB);                                                       \
    } while (0)

// Lowest set bit's index, cleared from m (64 when m is empty).
__device__ __forceinline__ int g3_pop(unsigned long long &m) {
    if (!m) return 64;
    const int b = __ffsll((long long)m) - 1;
    m &= m - 1;
    return b;
}

__global__ void __launch_bounds__(G3_THREADS)
k_step3_rows(TileOperands op, const int *__restrict__ Cptr, const int *__restrict__ Ccol, int tilem,
             const unsigned short *__restrict__ Cmask, const unsigned char *__restrict__ Cptr16,
             const int *__restrict__ Cnnz, const unsigned char *__restrict__ rowcls,
             unsigned char *__restrict__ Ccol8, double *__restrict__ Cval, G3Layout L) {
    extern __shared__ __align__(16) unsigned char g3_smem[];
    size_t off[G3_NARR];
    g3_layout(L, off);
    const int lamax = L.lamax, qb = L.qb, amax = L.amax, bnnzs = L.bnnz;
    double *s_acc = reinterpret_cast<double *>(g3_smem + off[0]);
    double *s_bval = reinterpret_cast<double *>(g3_smem + off[1]);
    double *s_aval = reinterpret_cast<double *>(g3_smem + off[2]);
    int *s_ccol = reinterpret_cast<int *>(g3_smem + off[1]);  // setup only: the B value area
    unsigned char *s_cidx = g3_smem + off[1] + 4 * G3_UMAX;    // setup only
    unsigned short *s_coff = reinterpret_cast<unsigned short *>(g3_smem + off[4]);  // < G3_NNZMAX
    unsigned char *s_pad = g3_smem + off[3];  // accumulator pads before each C tile
    unsigned short *s_cmask = reinterpret_cast<unsigned short *>(g3_smem + off[5]);
    unsigned char *s_cptr = g3_smem + off[6];
    int *s_row = reinterpret_cast<int *>(g3_smem + off[7]);
    unsigned char *s_pslot = g3_smem + off[8];
    unsigned short *s_bm = reinterpret_cast<unsigned short *>(g3_smem + off[9]);
    unsigned char *s_bp = g3_smem + off[10];
    unsigned short *s_bnnz = reinterpret_cast<unsigned short *>(g3_smem + off[11]);
    unsigned short *s_bnz = reinterpret_cast<unsigned short *>(g3_smem + off[12]);
    __shared__ unsigned short s_am[2][16];
    __shared__ unsigned char s_ap[2][16];
    __shared__ int s_wne[G3_THREADS / 64], s_wbig[G3_THREADS / 64];
    __shared__ unsigned long long s_stg, s_own[G3_THREADS / 64];  // staged steps; singles per wave
    static_assert(G3_UMAX <= G3_THREADS, "one C tile per thread in the setup");
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int I = (int)tsg_xcd_block(blockIdx.x, gridDim.x);
    if (I >= tilem || rowcls[I] != 1) return;  // block-uniform
    const int c0 = Cptr[I], u = Cptr[I + 1] - c0;
    const int nbase = Cnnz[c0], nrow = Cnnz[c0 + u] - nbase;
    const int a0 = op.Aptr[I], la = op.Aptr[I + 1] - a0;
    // C tiles of the tile-row (u <= G3_UMAX), one per thread.  Only the
    // nonempty ones get accumulator metadata (offset, masks, row offsets),
    // at their rank among the nonempty tiles: a task is active only when
    // one of its products exists, so no task lands in an empty C tile.
    // (their masks and row offsets are loaded here, stored after the barrier)
    int ccol = 0, cnz0 = 0, cnz1 = 0;
    bool ne = false;
    uint4 cmk0 = make_uint4(0u, 0u, 0u, 0u), cmk1 = cmk0, cpt = cmk0;
    if (tid < u) {
        ccol = Ccol[c0 + tid];
        cnz0 = Cnnz[c0 + tid];
        cnz1 = Cnnz[c0 + tid + 1];
        ne = cnz1 > cnz0;
        const uint4 *cm = reinterpret_cast<const uint4 *>(Cmask + (long long)(c0 + tid) * 16);
        cmk0 = cm[0];
        cmk1 = cm[1];
        cpt = reinterpret_cast<const uint4 *>(Cptr16 + (long long)(c0 + tid) * 16)[0];
    }
    const unsigned long long neb = __ballot(ne), bgb = __ballot(cnz1 - cnz0 >= G3_PADMIN);
    if (lane == 0) {
        s_wne[wave] = __popcll(neb);
        s_wbig[wave] = __popcll(bgb);
    }
    // row table: one entry per A tile of the tile-row (wave 0, la <= 64)
    if (wave == 0) {
        int lb = 0, pan = 0, rk = 0;
        if (lane < la) {
            const int p = a0 + lane, K = op.Acol[p];
            const int pa0 = op.Annz[p], q0 = op.Bptr[K], q1 = op.Bptr[K + 1];
            pan = op.Annz[p + 1] - pa0;
            const int pb0 = op.Bnnz[q0], pb1 = op.Bnnz[q1];
            // the row masks: a single entry's row and column
            const uint4 *amk = reinterpret_cast<const uint4 *>(op.Amask + (long long)p * 16);
            const uint4 m0 = amk[0], m1 = amk[1];
            const unsigned w[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
#pragma unroll
            for (int r = 15; r >= 0; r--) {
                const unsigned m = (w[r >> 1] >> (16 * (r & 1))) & 0xffffu;
                if (m) rk = r | ((__clz((int)m) - 16) << 8);
            }
            lb = q1 - q0;
            SROW(1, lane) = pa0;
            SROW(3, lane) = q0;
            SROW(4, lane) = pan == 1 ? rk : pb0;
            SROW(5, lane) = pb1 - pb0;
        }
        const int incl = tsg_wave_incl_scan(lb);
        if (lane < la) SROW(2, lane) = pan | (lb << 9) | ((incl - lb) << 15);
        const unsigned long long stg = __ballot(pan > 1);
        unsigned long long ow[G3_THREADS / 64];
#pragma unroll
        for (int w = 0; w < G3_THREADS / 64; w++) ow[w] = __ballot(pan == 1 && g3_owner(rk & 15) == w);
        if (lane == 0) {
            s_stg = stg;
#pragma unroll
            for (int w = 0; w < G3_THREADS / 64; w++) s_own[w] = ow[w];
        }
    }
    __syncthreads();
    // the ring's first three stages need only the row table: their loads
    // are in flight while the C tile slots and pair slots are set up
    unsigned long long rest = s_stg, own = s_own[wave];
    int fa = g3_pop(rest), fb = g3_pop(rest);  // staged steps 0 and 1 (64: none)
    G3_SLOT_DECL(R0_);
    G3_SLOT_DECL(R1_);
    if (fa < 64) G3_LOAD(fa, R0_);
    if (fb < 64) G3_LOAD(fb, R1_);
    unsigned sm_m = 0u, sm_n = 0u, sm_p = 0u, sv_m = 0u, sv_cm = 0u;
    int sv_p = 0, sv_cb = 0;
    double sv_0 = 0.0, sv_1 = 0.0, sv_a = 0.0;
    int sA = g3_pop(own), sB = g3_pop(own);  // this wave's next two single-entry steps
    if (sA < 64) G3_L1(sA);
    int ci = __popcll(neb & ((1ull << lane) - 1ull)), pad = __popcll(bgb & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; w++) {
        ci += s_wne[w];
        pad += s_wbig[w];
    }
    int un = 0, npad = 0;
#pragma unroll
    for (int w = 0; w < G3_THREADS / 64; w++) {
        un += s_wne[w];
        npad += s_wbig[w];
    }
    // C tile slot of every (A tile, B tile) pair of the tile-row.  When the
    // row's C tile columns span at most the accumulator's bytes, the
    // accumulator area first holds a direct map column -> slot (every pair's
    // column is in the row's C tile list, so unmapped bytes are never read);
    // otherwise a binary search over the sorted C tile columns.  An empty
    // tile maps to the next nonempty tile's slot (< 256); no task uses it.
    const int jlo = Ccol[c0], jspan = Ccol[c0 + u - 1] - jlo + 1;
    const bool direct = jspan <= 8 * L.nnzmax;
    unsigned char *s_map = reinterpret_cast<unsigned char *>(s_acc);
    if (tid < u) {
        s_ccol[tid] = ccol;
        s_cidx[tid] = (unsigned char)ci;
        if (direct) s_map[ccol - jlo] = (unsigned char)ci;
        if (ne) {
            s_coff[ci] = (unsigned short)(cnz0 - nbase);
            s_pad[ci] = (unsigned char)pad;
            reinterpret_cast<uint4 *>(s_cmask)[2 * ci] = cmk0;
            reinterpret_cast<uint4 *>(s_cmask)[2 * ci + 1] = cmk1;
            reinterpret_cast<uint4 *>(s_cptr)[ci] = cpt;
        }
    }
    // B tile columns of the row's (A tile j, B tile qi) pairs, xx = j * G3_QB
    // + qi: those of the first 32 A tiles are loaded here, so that their
    // latency overlaps the barrier (the rest after it)
    static_assert(G3_QB == 32, "pair index split by shifts");
    static_assert(G3_LAMAX * G3_QB <= 8 * G3_THREADS, "eight pairs per thread");
    int pj[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int xx = tid + G3_THREADS * t, j = xx >> 5, qi = xx & 31;
        pj[t] = xx < la * G3_QB && qi < G3_LB(j) ? op.Bcol[SROW(3, j) + qi] : -1;
    }
    __syncthreads();
    // the pair's slot index is its rank G3_PB(j) + qi
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int xx = tid + G3_THREADS * t, j = xx >> 5, qi = xx & 31;
        const int J = t < 4 ? pj[t & 3] : xx < la * G3_QB && qi < G3_LB(j) ? op.Bcol[SROW(3, j) + qi] : -1;
        if (J < 0) continue;
        const int x = G3_PB(j) + qi;
        if (direct) {
            s_pslot[x] = s_map[J - jlo];
        } else {
            int lo = 0, hi = u - 1;  // J is in the step-1 list
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_ccol[mid] < J) lo = mid + 1; else hi = mid;
            }
            s_pslot[x] = s_cidx[lo];
        }
    }
    __syncthreads();  // the map area becomes the accumulator
    for (int i = tid; i < nrow + npad; i += G3_THREADS) s_acc[i] = 0.0;
    if (sA < 64) G3_L2(sA);
    if (sB < 64) G3_L1(sB);
    if (fa < 64) G3_STORE(fa, R0_, 0);
    __syncthreads();

    // compute staged K-step j from buffer buf: the tasks of this wave's rows
    auto compute = [&](int j, int buf) {
        const int nq = G3_LB(j), pbase = G3_PB(j);
        const int ncand = nq * 4;
        const unsigned short *sbz = s_bnz + buf * qb;
        // candidate tasks (B tile qi, A row rb + 4i) meeting at least one
        // nonempty B row, compacted into a register queue: position Q of the
        // queue is lane Q & 63 of qlo (Q < 64) or qhi (Q >= 64).  There are
        // at most two chunks of 64 candidates (ncand <= G3_QB * 4 = 128), so
        // qn <= 64 before each chunk.  Each chunk is moved by one forward
        // permute over all 64 lanes: active lanes to the queue positions
        // qn.., inactive ones after them, so every lane receives exactly one
        // value and the permute needs no LDS memory.
        int qn = 0, qlo = 0, qhi = 0;
        const int rb = g3_rbase(wave);
        for (int c = 0; c < ncand; c += 64) {
            const int x = c + lane;
            int code = 0;
            bool act = false;
            if (x < ncand) {
                const int qi = x >> 2, rr = rb + 4 * (x & 3);
                act = (s_am[buf][rr] & sbz[qi]) != 0;
                code = (qi << 4) | rr;
            }
            const unsigned long long bal = __ballot(act);
            const int na = __popcll(bal), below = __popcll(bal & ((1ull << lane) - 1ull));
            const int rank = act ? below : na + (lane - below);
            const int got = __builtin_amdgcn_ds_permute(((qn + rank) & 63) << 2, code);
            if (lane >= qn && lane < qn + na) qlo = got;
            if (lane + 64 < qn + na) qhi = got;
            qn += na;
        }
        const unsigned short *sbm = s_bm + buf * qb * 16;
        const unsigned char *sbp = s_bp + buf * qb * 16;
        const unsigned short *sbnnz = s_bnnz + buf * (qb + 1);
        const double *sbv = s_bval + buf * bnnzs;
        const double *sav = s_aval + buf * amax;
        for (int t0 = 0; t0 < qn; t0 += 64) {
            if (t0 + lane >= qn) break;
            const int code = t0 ? qhi : qlo;
            const int qi = code >> 4, rr = code & 15;
            const int slot = s_pslot[pbase + qi];
            const unsigned cm = s_cmask[slot * 16 + rr];
            const int cbase = s_coff[slot] + s_pad[slot] + s_cptr[slot * 16 + rr];
            const unsigned short *bm16 = sbm + qi * 16;
            const unsigned char *bp16 = sbp + qi * 16;
            const double *bvt = sbv + sbnnz[qi];
            unsigned am = s_am[buf][rr];
            const double *av = sav + s_ap[buf][rr];
            int ia = 0;
            while (am) {  // A row entries, ascending k
                const int b = 31 - __clz(am);
                am &= ~(1u << b);
                const int k = 15 - b;
                const double a = av[ia++];
                unsigned bm = bm16[k];
                const double *brow = bvt + bp16[k];
                int jb = 0;
                while (bm) {  // B row k entries, ascending column
                    const int bb = 31 - __clz(bm);
                    bm &= ~(1u << bb);
                    double *acc = &s_acc[cbase + __popc(cm >> (bb + 1))];
                    *acc = __builtin_fma(a, brow[jb++], *acc);
                }
            }
        }
    };
    // single-entry steps before the first staged step
    while (sA < fa) G3_SINGLE(sA);
    // staged step fa (number f) from buffer f & 1 and ring slot SR; then the
    // next staged step fb is stored from slot SN, the singles between them
    // are taken, and staged step fc is loaded into slot SR.  A barrier
    // precedes every staged step: its stores are visible, and the buffer it
    // overwrote was last read two staged steps before.
    int f = 0;
#define G3_ITER(SN, SR)                                 \
    if (fa < 64) {                                      \
        compute(fa, f & 1);                             \
        if (fb < 64) G3_STORE(fb, SN, (f + 1) & 1);     \
        const int fc = g3_pop(rest);                    \
        while (sA < fb) G3_SINGLE(sA);                  \
        if (fc < 64) G3_LOAD(fc, SR);                   \
        if (fb < 64) __syncthreads();                   \
        fa = fb;                                        \
        fb = fc;                                        \
        f++;                                            \
    }
    while (fa < 64) {
        G3_ITER(R1_, R0_)
        G3_ITER(R0_, R1_)
    }
#undef G3_ITER
    __syncthreads();  // every wave's rows are complete
    // write out, one C tile row at a time: values from the padded
    // accumulator, columns from the masks
    for (int x = tid; x < un * 16; x += G3_THREADS) {
        const int slot = x >> 4;
        unsigned cm = s_cmask[x];
        const int fo = s_coff[slot] + s_cptr[x];
        const double *av = s_acc + fo + s_pad[slot];
        unsigned char *oc = Ccol8 + nbase + fo;
        double *ov = Cval + (long long)nbase + fo;
        int jj = 0;
        while (cm) {
            const int bb = 31 - __clz(cm);
            cm &= ~(1u << bb);
            oc[jj] = (unsigned char)(15 - bb);
            ov[jj] = av[jj];
            jj++;
        }
    }
}
#undef SROW

// Step-3 class of every C tile-row: 1 = row-wise kernel (its C tiles, C
// nonzeros, A tiles and B tile-rows are within the G3 caps), 0 = tile-parallel
// kernel, 2 = nothing to compute.  Per block, part[8 * block + i]: i = 0 the
// class-0 rows, i = 1..6 the maxima
// over class-1 rows of (nonempty C tiles, C nonzeros plus accumulator pads,
// A tiles, B tiles per B tile-row, B nonzeros per B tile-row, A tile
// nonzeros) for the LDS layout.
// 16 lanes per tile-row (16 tile-rows per block): the lanes walk the row's A
// tiles and C tiles in strides of 16, so the dependent loads of different A
// tiles are in flight together.  k_row_class_reduce folds the blocks' partials
// into out[0] (sum) and out[1..6] (maxima): no atomics, which on one cache
// line would serialise across the thousands of blocks.
__device__ __forceinline__ int g16_max(int v) {
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 16));
    return v;
}
__global__ void __launch_bounds__(256)
k_row_class(TileOperands op, const int *__restrict__ Cptr, const int *__restrict__ Cnnz, int tilem,
            unsigned char *__restrict__ cls, int *__restrict__ part, int none_fit) {
    __shared__ int s_red[4][7];
    const int r = threadIdx.x & 15, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = blockIdx.x * 16 + (threadIdx.x >> 4);
    int unfit = 0, mu = 0, mn = 0, ml = 0, mq = 0, mb = 0, ma = 0;
    if (I < tilem) {  // uniform in the 16-lane group
        const int c0 = Cptr[I], c1 = Cptr[I + 1];
        const int nrow = Cnnz[c1] - Cnnz[c0];
        unsigned char k;
        if (nrow == 0) {
            k = 2;
        } else {
            const int a0 = op.Aptr[I], a1 = op.Aptr[I + 1];
            bool ok = !none_fit && c1 - c0 <= G3_UMAX && nrow <= G3_NNZMAX && a1 - a0 <= G3_LAMAX;
            if (ok) {  // group-uniform
                int bad = 0, lq = 0, lbn = 0, lan = 0, ne = 0;
                for (int p = a0 + r; p < a1; p += 16) {
                    const int K = op.Acol[p];
                    const int q0 = op.Bptr[K], q1 = op.Bptr[K + 1];
                    const int bn = op.Bnnz[q1] - op.Bnnz[q0];
                    bad |= q1 - q0 > G3_QB || bn > G3_BNNZ;
                    lq = max(lq, q1 - q0);
                    lbn = max(lbn, bn);
                    lan = max(lan, op.Annz[p + 1] - op.Annz[p]);
                }
                // nonempty C tiles (the row-wise kernel's slots) and those
                // followed by an accumulator pad
                int nbig = 0;
                for (int t = c0 + r; t < c1; t += 16) {
                    const int tn = Cnnz[t + 1] - Cnnz[t];
                    ne += tn > 0;
                    nbig += tn >= G3_PADMIN;
                }
                ok = g16_max(bad) == 0;
                ne = tsg_group16_sum(ne);
                nbig = tsg_group16_sum(nbig);
                lq = g16_max(lq); lbn = g16_max(lbn); lan = g16_max(lan);
                if (ok) {
                    mu = ne; mn = nrow + nbig; ml = a1 - a0; mq = lq; mb = lbn; ma = lan;
                }
            }
            k = ok ? 1 : 0;
        }
        if (r == 0) cls[I] = k;
        unfit = r == 0 && k == 0;
    }
    unfit = tsg_wave_sum(unfit);
    mu = tsg_wave_max(mu); mn = tsg_wave_max(mn); ml = tsg_wave_max(ml);
    mq = tsg_wave_max(mq); mb = tsg_wave_max(mb); ma = tsg_wave_max(ma);
    if (lane == 0) {
        s_red[wave][0] = unfit; s_red[wave][1] = mu; s_red[wave][2] = mn; s_red[wave][3] = ml;
        s_red[wave][4] = mq; s_red[wave][5] = mb; s_red[wave][6] = ma;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int i = threadIdx.x;
        int v = 0;
        if (i < 7) {
            v = s_red[0][i];
#pragma unroll
            for (int w = 1; w < 4; w++) v = i == 0 ? v + s_red[w][i] : max(v, s_red[w][i]);
        }
        part[8 * (long long)blockIdx.x + i] = v;
    }
}

// out[0] = sum and out[1..6] = maxima of the nb blocks' k_row_class partials
// (one block of 1024 threads); out[7] = 0.
__global__ void __launch_bounds__(1024)
k_row_class_reduce(const int *__restrict__ part, int nb, int *__restrict__ out) {
    __shared__ int s_r[16][7];
    int v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 1024) {
        v[0] += part[8 * (long long)b];
#pragma unroll
        for (int i = 1; i < 7; i++) v[i] = max(v[i], part[8 * (long long)b + i]);
    }
    v[0] = tsg_wave_sum(v[0]);
#pragma unroll
    for (int i = 1; i < 7; i++) v[i] = tsg_wave_max(v[i]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 7; i++) s_r[wave][i] = v[i];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int i = threadIdx.x;
        int r = 0;
        if (i < 7) {
            r = s_r[0][i];
            for (int w = 1; w < 16; w++) r = i == 0 ? r + s_r[w][i] : max(r, s_r[w][i]);
        }
        out[i] = r;
    }
}

// ------------------------------------------------------------------ host
// Step-1 count only: number of C tiles of A*B (structurally empty ones
// included), from the row-major tile indices.
extern "C" int tsg_spgemm_count_tiles(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, long long *out) {
    if (A->n != B->m) { tsg_set_error("count_tiles: inner dimensions differ"); return TSG_ERR_ARG; }
    int *cnt;
    TSG_TRY(tsg_alloc_t(ctx, &cnt, (size_t)A->tilem + 1));
    if (A->tilem > 0) {
        hipLaunchKernelGGL((k_step1<false, 256>), dim3(tsg_div_up(A->tilem, S1_WAVES)), dim3(64 * S1_WAVES), 0,
                           ctx->stream, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr, B->tile_colidx,
                           (const int4 *)nullptr, (const unsigned *)nullptr, (const int *)nullptr, cnt,
                           (int *)nullptr, (int *)nullptr, (unsigned *)nullptr, (int *)nullptr);
        TSG_LAUNCH_CHECK();
    }
    long long tot = 0;
    int st = tsg_scan_i32(ctx, cnt, cnt, A->tilem, &tot);
    tsg_free(ctx, cnt);
    *out = tot;
    return st;
}

// ------------------------------------------------------------------ no-allocation tile count
// tsg_spgemm_count_tiles_noalloc keeps its scratch in one fixed device
// array per GPU (loaded with the code object, held until the process
// ends): the per-tile-row C tile counts, then the per-block partial sums.
// Host code never passes the array's name to a kernel (there it is only a
// host-side stand-in); the device address comes from hipGetSymbolAddress,
// taken once per device with that device current, and kernels receive it
// as an argument.
#define NOALLOC_PARTS 1024  // partial sums (int64), at the end of the array
static_assert(TSG_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch is a whole number of 256-byte lines");
static_assert((long long)TSG_NOALLOC_MAX_TILE_ROWS * 4 + NOALLOC_PARTS * 8 <= TSG_NOALLOC_SCRATCH_BYTES,
              "counts and partials fit the scratch array");
static __device__ __align__(256) unsigned char scratch[TSG_NOALLOC_SCRATCH_BYTES];

#define NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

static int noalloc_scratch(int device, unsigned char **out) {
    *out = nullptr;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        tsg_set_error("count_tiles_noalloc: device %d outside [0, %d)", device, NOALLOC_MAX_DEVICES);
        return TSG_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        int cur = -1;
        TSG_HIP(hipGetDevice(&cur));
        if (cur != device) TSG_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            tsg_set_error("count_tiles_noalloc: hipGetSymbolAddress(scratch) on device %d -> %s", device,
                          hipGetErrorString(e));
            return TSG_ERR_HIP;
        }
        g_scratch_dev[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[device];
    return TSG_OK;
}

// Sum of n int32 counts: block b adds a grid-stride share into part[b]
// (every slot written, so nothing needs clearing), then one block adds the
// partials into *out.
__global__ void __launch_bounds__(256)
k_noalloc_partial(const int *__restrict__ cnt, int n, long long *__restrict__ part) {
    long long s = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += cnt[i];
    s = tsg_wave_sum(s);
    __shared__ long long w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}
__global__ void __launch_bounds__(256)
k_noalloc_final(const long long *__restrict__ part, int nparts, long long *__restrict__ out) {
    long long s = 0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    s = tsg_wave_sum(s);
    __shared__ long long w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = w[0] + w[1] + w[2] + w[3];
}

extern "C" int tsg_spgemm_count_tiles_noalloc(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B,
                                              long long *d_out) {
    if (!ctx || !A || !B || !d_out) { tsg_set_error("count_tiles_noalloc: null argument"); return TSG_ERR_ARG; }
    if (A->n != B->m) { tsg_set_error("count_tiles_noalloc: inner dimensions differ"); return TSG_ERR_ARG; }
    if (A->tilem < 0 || A->tilem > TSG_NOALLOC_MAX_TILE_ROWS) {
        tsg_set_error("count_tiles_noalloc: A has %d tile-rows; the fixed scratch holds at most %d", A->tilem,
                      TSG_NOALLOC_MAX_TILE_ROWS);
        return TSG_ERR_ARG;
    }
    if (A->tilem > 0 && (!A->tile_ptr || !A->tile_colidx || !B->tile_ptr || !B->tile_colidx)) {
        tsg_set_error("count_tiles_noalloc: missing tile index arrays");
        return TSG_ERR_ARG;
    }
    unsigned char *scr;
    TSG_TRY(noalloc_scratch(ctx->device, &scr));
    int *cnt = reinterpret_cast<int *>(scr);
    long long *part = reinterpret_cast<long long *>(scr + TSG_NOALLOC_SCRATCH_BYTES - NOALLOC_PARTS * 8);
    int nparts = (int)std::min<long long>(NOALLOC_PARTS, std::max(1u, tsg_div_up(A->tilem, 256)));
    if (A->tilem > 0) {
        hipLaunchKernelGGL((k_step1<false, 256>), dim3(tsg_div_up(A->tilem, S1_WAVES)), dim3(64 * S1_WAVES), 0,
                           ctx->stream, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr, B->tile_colidx,
                           (const int4 *)nullptr, (const unsigned *)nullptr, (const int *)nullptr, cnt,
                           (int *)nullptr, (int *)nullptr, (unsigned *)nullptr, (int *)nullptr);
        TSG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_noalloc_partial, dim3(nparts), dim3(256), 0, ctx->stream, (const int *)cnt, A->tilem, part);
    hipLaunchKernelGGL(k_noalloc_final, dim3(1), dim3(256), 0, ctx->stream, (const long long *)part, nparts, d_out);
    TSG_LAUNCH_CHECK();
    return TSG_OK;
}

static int ev_ms(hipEvent_t a, hipEvent_t b, double *ms) {
    float f = 0.f;
    TSG_HIP(hipEventElapsedTime(&f, a, b));
    *ms = f;
    return TSG_OK;
}

// Launch helpers shared by the synchronous pipeline and the enqueue-only
// (graph-capturable) one.
static TileOperands make_operands(const tsg_dtiles *A, const tsg_dtiles *B) {
    TileOperands op;
    op.Aptr = A->tile_ptr; op.Acol = A->tile_colidx; op.Annz = A->tile_nnz; op.Aptr16 = A->csr_ptr;
    op.Amask = A->mask; op.Aval = A->val;
    op.Bptr = B->tile_ptr; op.Bcol = B->tile_colidx;
    op.Bcscptr = B->csc_ptr; op.Bcscrow = B->csc_rowidx; op.Bcscid = B->csc_tileid; op.Bnnz = B->tile_nnz;
    op.Bptr16 = B->csr_ptr; op.Bmask = B->mask; op.Bval = B->val;
    op.Arowp = A->rowx_ptr; op.Acolx = A->rowx_col; op.Avalx = A->rowx_val;
    return op;
}

// Step-1 bitmap window in 32-bit words per wave: 256 (8192 tile columns,
// 8 KiB of LDS per block), 1024 (32768, 16 KiB) or 2048 (65536, 32 KiB).
// By default B with 8193..32768 tile columns gets the 1024-word window and
// B with 32769..65536 the 2048-word one, so a row of a product without band
// structure needs one window, with no more LDS (and, for the kept-row
// bitmap ORs, registers) than that takes; wider B (banded operands such as
// stencils) keeps the narrow one.  The count pass's bitmaps are kept for
// the fill pass up to 1 GiB.
int tsg_s1_words(const tsg_dtiles *B, const tsg_tune &t) {
    if (t.s1_window_words == 256 || t.s1_window_words == 1024 || t.s1_window_words == 2048) return t.s1_window_words;
    if (B->tilen > 8192 && B->tilen <= 32768) return 1024;
    return B->tilen > 32768 && B->tilen <= 65536 ? 2048 : 256;
}
static size_t keep_bytes(long long tilem, int words) {
    const size_t b = (size_t)tilem * (size_t)words * 4;
    return b <= ((size_t)1 << 30) ? b : 0;
}

// B's tile-row bitmaps for step 1 (k_brow_desc, k_brow_fill): a descriptor
// per B tile-row, then the words.  0 when B has no tile row index or the
// word offsets would not fit 32 bits (step 1 then enumerates every
// candidate).
static size_t brow_bits_bytes(const tsg_dtiles *B) {
    if (!B || B->tilem <= 0 || !B->tile_rowidx) return 0;
    const long long words = 2ll * B->numtile + 2ll * B->tilem;
    if (words >= 0x7fffffffLL) return 0;
    return (size_t)B->tilem * 16 + (size_t)words * 4;
}
static int launch_brow_bits(hipStream_t s, const tsg_dtiles *B, void *buf, const int4 **desc,
                            const unsigned **bits, int *any) {
    *desc = nullptr;
    *bits = nullptr;
    if (!buf || !brow_bits_bytes(B)) return TSG_OK;
    int4 *d = (int4 *)buf;
    unsigned *w = (unsigned *)((char *)buf + (size_t)B->tilem * 16);
    hipLaunchKernelGGL(k_brow_desc, dim3(tsg_div_up(B->tilem, 256)), dim3(256), 0, s, (const int *)B->tile_ptr,
                       (const int *)B->tile_colidx, B->tilem, d, any);
    if (B->numtile > 0)
        hipLaunchKernelGGL(k_brow_fill, dim3(tsg_div_up(B->numtile, 256)), dim3(256), 0, s,
                           (const int *)B->tile_rowidx, (const int *)B->tile_colidx, (long long)B->numtile,
                           (const int4 *)d, w, (const int *)any);
    TSG_LAUNCH_CHECK();
    *desc = d;
    *bits = w;
    return TSG_OK;
}

static int launch_step1(hipStream_t s, const tsg_dtiles *A, const tsg_dtiles *B, tsg_dtiles *C, bool fill,
                        int words, unsigned *bmstore, const int4 *bdesc, const unsigned *bbits,
                        const int *bany, int *rowspan = nullptr) {
    if (C->tilem <= 0) return TSG_OK;
    const dim3 g(tsg_div_up(C->tilem, S1_WAVES)), blk(64 * S1_WAVES);
    const bool wide = words == 2048;
    if (words == 1024) {
        if (!fill)
            hipLaunchKernelGGL((k_step1<false, 1024>), g, blk, 0, s, A->tile_ptr, A->tile_colidx, A->tilem,
                               B->tile_ptr, B->tile_colidx, bdesc, bbits, bany, C->tile_ptr, (int *)nullptr,
                               (int *)nullptr, bmstore, rowspan);
        else
            hipLaunchKernelGGL((k_step1<true, 1024>), g, blk, 0, s, A->tile_ptr, A->tile_colidx, A->tilem,
                               B->tile_ptr, B->tile_colidx, bdesc, bbits, bany, C->tile_ptr, C->tile_colidx,
                               C->tile_rowidx, bmstore, (int *)nullptr);
    } else if (!fill && !wide)
        hipLaunchKernelGGL((k_step1<false, 256>), g, blk, 0, s, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr,
                           B->tile_colidx, bdesc, bbits, bany, C->tile_ptr, (int *)nullptr, (int *)nullptr, bmstore, rowspan);
    else if (!fill)
        hipLaunchKernelGGL((k_step1<false, 2048>), g, blk, 0, s, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr,
                           B->tile_colidx, bdesc, bbits, bany, C->tile_ptr, (int *)nullptr, (int *)nullptr, bmstore, rowspan);
    else if (!wide)
        hipLaunchKernelGGL((k_step1<true, 256>), g, blk, 0, s, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr,
                           B->tile_colidx, bdesc, bbits, bany, C->tile_ptr, C->tile_colidx, C->tile_rowidx, bmstore,
                           (int *)nullptr);
    else
        hipLaunchKernelGGL((k_step1<true, 2048>), g, blk, 0, s, A->tile_ptr, A->tile_colidx, A->tilem, B->tile_ptr,
                           B->tile_colidx, bdesc, bbits, bany, C->tile_ptr, C->tile_colidx, C->tile_rowidx, bmstore,
                           (int *)nullptr);
    TSG_LAUNCH_CHECK();
    return TSG_OK;
}

// Device scratch of one SpGEMM beyond C itself.
struct SpgemmWork {
    long long *part1;  // scan partials: C tile pointers
    long long *part2;  // scan partials: C tile nnz
    long long *part3;  // scan partials: segment counts
    int *segcnt;       // tilem+1 segment counts -> offsets (total at [tilem])
    unsigned char *rowcls;
    int *cls_out;
    int *cls_part;  // k_row_class partials, 8 ints per block
    void *bbits;       // B's tile-row bitmaps for step 1 (may be NULL)
    int4 *seg2;        // step-2 segments
    unsigned *bmstore; // step-1 bitmaps kept from the count pass (may be NULL)
    int4 *seg3;        // step-3 element segments
    long long seg2_cap, seg3_cap;
    // heavy rows (synchronous pipeline only; NULL = none)
    unsigned char *hflag;  // per C row
    int2 *hitems;          // heavy-row work items
    int *hcount;           // their number (device)
    int nheavy, nitems, hwaves;
    int *cursor;           // hwaves cursor arrays of hstride ints
    long long hstride;
    int *hnext;            // work-item counter (zeroed before the launch)
};

static size_t up16(size_t x) { return (x + 15) & ~(size_t)15; }

// Offsets of the work arrays inside one buffer; returns its size.  The
// step-3 segments go last (the synchronous pipeline adds them only once
// C's nonzero count is known).
#define WORK_NARR 11

static size_t work_layout(long long tilem, long long numtile, long long nnz, size_t off[WORK_NARR],
                          int s1w = 256, size_t bbytes = 0) {
    const size_t sz[WORK_NARR] = {
        8 * (size_t)tsg_scan_partial_count(tilem), 8 * (size_t)tsg_scan_partial_count(numtile),
        8 * (size_t)tsg_scan_partial_count(tilem), 4 * ((size_t)tilem + 1), (size_t)tilem + 1, 32,
        16 * (size_t)seg_bound(tilem, numtile, G2_UMAX, 0, 0),
        keep_bytes(tilem, s1w),  // step-1 bitmaps (enqueue-only pipeline)
        32 * (size_t)((tilem + 15) / 16 + 1),  // row-class partials, 8 ints per block
        bbytes,                                // B's tile-row bitmaps (enqueue-only pipeline)
        16 * (size_t)(tilem + nnz / (G3E_NNZ - 255) + 2)};
    size_t o = 0;
    for (int i = 0; i < WORK_NARR; i++) {
        off[i] = o;
        o += up16(sz[i]);
    }
    return o;
}

static SpgemmWork work_at(void *base, long long tilem, long long numtile, long long nnz, void *seg3_base,
                          int s1w = 256, size_t bbytes = 0) {
    size_t off[WORK_NARR];
    work_layout(tilem, numtile, nnz, off, s1w, bbytes);
    char *b = (char *)base;
    SpgemmWork w;
    w.part1 = (long long *)(b + off[0]);
    w.part2 = (long long *)(b + off[1]);
    w.part3 = (long long *)(b + off[2]);
    w.segcnt = (int *)(b + off[3]);
    w.rowcls = (unsigned char *)(b + off[4]);
    w.cls_out = (int *)(b + off[5]);
    w.seg2 = (int4 *)(b + off[6]);
    w.bmstore = keep_bytes(tilem, s1w) ? (unsigned *)(b + off[7]) : nullptr;
    w.cls_part = (int *)(b + off[8]);
    w.bbits = bbytes ? (void *)(b + off[9]) : nullptr;
    w.seg3 = (int4 *)(seg3_base ? (char *)seg3_base : b + off[10]);
    w.seg2_cap = seg_bound(tilem, numtile, G2_UMAX, 0, 0);
    w.seg3_cap = tilem + nnz / (G3E_NNZ - 255) + 2;
    w.hflag = nullptr;
    w.hitems = nullptr;
    w.hcount = nullptr;
    w.nheavy = w.nitems = w.hwaves = 0;
    w.cursor = nullptr;
    w.hstride = 0;
    w.hnext = nullptr;
    return w;
}

// Segments of the selected tile-rows into items; the count lands in
// w.segcnt[tilem].
template <int MODE>
static int launch_segments(hipStream_t s, const tsg_dtiles *C, const SpgemmWork &w, int want, int max_tiles,
                           int max_nnz, int4 *items, const unsigned char *cls, long long cap = 0) {
    const unsigned g = tsg_div_up(C->tilem, 256);
    hipLaunchKernelGGL(k_seg<MODE>, dim3(g), dim3(256), 0, s, C->tile_ptr, C->tile_nnz, C->tilem, cls, want,
                       max_tiles, max_nnz, w.segcnt, (int4 *)nullptr);
    TSG_LAUNCH_CHECK();
    TSG_TRY(tsg_scan_i32_enqueue(s, w.segcnt, w.segcnt, C->tilem, w.part3));
    if (MODE == 1) {  // one thread per segment, cap = the segment array's capacity (>= the count)
        if (cap > 0) {
            hipLaunchKernelGGL(k_seg1_fill, dim3((unsigned)tsg_div_up(cap, 256)), dim3(256), 0, s, C->tile_ptr,
                               C->tile_nnz, C->tilem, (const int *)w.segcnt, max_nnz, items);
            TSG_LAUNCH_CHECK();
        }
        return TSG_OK;
    }
    hipLaunchKernelGGL(k_seg<MODE>, dim3(g), dim3(256), 0, s, C->tile_ptr, C->tile_nnz, C->tilem, cls, want,
                       max_tiles, max_nnz, w.segcnt, items);
    TSG_LAUNCH_CHECK();
    return TSG_OK;
}

static int launch_step2(hipStream_t s, const TileOperands &op, const tsg_dtiles *B, tsg_dtiles *C,
                        const tsg_spgemm_plan &P, const SpgemmWork &w) {
    if (C->numtile <= 0) return TSG_OK;
    if (P.all) {
        const unsigned nblk = tp_grid(tsg_div_up(C->numtile, S23_WAVES * TILES_PER_WAVE));
        hipLaunchKernelGGL(k_step2_tiles, dim3(nblk), dim3(64 * S23_WAVES), 0, s, op, C->tile_ptr, C->tile_rowidx,
                           C->tile_colidx, C->numtile, C->mask, C->csr_ptr, C->tile_nnz, 1);
        TSG_LAUNCH_CHECK();
        return TSG_OK;
    }
    if (P.s3_elem) {  // B has its row index: element-wise symbolic for hypersparse tile-rows
        hipLaunchKernelGGL(k_step2_kind, dim3(tsg_div_up(C->tilem, 256)), dim3(256), 0, s, op.Aptr, op.Annz,
                           C->tilem, P.elem_all, w.rowcls);
        TSG_LAUNCH_CHECK();
    }
    TSG_TRY(launch_segments<0>(s, C, w, -1, G2_UMAX, 0, w.seg2, P.s3_elem ? w.rowcls : nullptr));
    const int s2u = P.s2u > 0 ? P.s2u : G2_UMAX;
    const size_t lds2 = (4 * G2_MSTRIDE + 4) * (size_t)s2u + (((size_t)P.s2map + 15) & ~(size_t)15);
    hipLaunchKernelGGL(k_step2_rows, dim3((unsigned)w.seg2_cap), dim3(G2_THREADS), lds2, s, op, (const int4 *)w.seg2,
                       (const int *)(w.segcnt + C->tilem), C->tile_ptr, C->tile_colidx, C->mask, C->csr_ptr,
                       C->tile_nnz, s2u, P.s2map);
    TSG_LAUNCH_CHECK();
    if (P.s3_elem) {
        hipLaunchKernelGGL(k_step2_elem, dim3((unsigned)w.seg2_cap), dim3(G2_THREADS), 0, s, op, (const int *)B->rowx_ptr,
                           (const int *)B->rowx_col, (const int4 *)w.seg2, (const int *)(w.segcnt + C->tilem),
                           C->tile_ptr, C->tile_colidx, C->mask, C->csr_ptr, C->tile_nnz);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

static int launch_row_class(hipStream_t s, const TileOperands &op, tsg_dtiles *C, unsigned char *rowcls,
                            int *cls_part, int *cls_out, int none_fit) {
    const int nb = (int)tsg_div_up(C->tilem, 16);
    hipLaunchKernelGGL(k_row_class, dim3(nb), dim3(256), 0, s, op, C->tile_ptr, C->tile_nnz, C->tilem, rowcls,
                       cls_part, none_fit);
    TSG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_row_class_reduce, dim3(1), dim3(1024), 0, s, (const int *)cls_part, nb, cls_out);
    TSG_LAUNCH_CHECK();
    return TSG_OK;
}

static int launch_step3(hipStream_t s, const TileOperands &op, const tsg_dtiles *B, tsg_dtiles *C,
                        const tsg_spgemm_plan &P, const SpgemmWork &w, hipEvent_t *ev = nullptr) {
    if (C->nnz <= 0) return TSG_OK;
    if (!P.all && P.s3_rows) {
        G3Layout L;
        L.umax = P.g3[0]; L.nnzmax = P.g3[1]; L.lamax = P.g3[2];
        L.qb = P.g3[3]; L.bnnz = P.g3[4]; L.amax = P.g3[5];
        size_t off[G3_NARR];
        const size_t lds = g3_layout(L, off);
        if (ev) TSG_HIP(hipEventRecord(ev[0], s));
        hipLaunchKernelGGL(k_step3_rows, dim3(C->tilem), dim3(G3_THREADS), lds, s, op, C->tile_ptr, C->tile_colidx,
                           C->tilem, C->mask, C->csr_ptr, C->tile_nnz, (const unsigned char *)w.rowcls, C->csr_col,
                           C->val, L);
        TSG_LAUNCH_CHECK();
        if (ev) TSG_HIP(hipEventRecord(ev[1], s));
    }
    if (!P.all && P.s3_unfit && P.s3_elem) {
        TSG_TRY(launch_segments<1>(s, C, w, 0, 0x7fffffff, G3E_NNZ, w.seg3, w.rowcls, w.seg3_cap));
        const bool rx = op.Arowp != nullptr;
        hipLaunchKernelGGL(P.s3pk ? (rx ? k_step3_elem<true, true> : k_step3_elem<true, false>)
                                  : (rx ? k_step3_elem<false, true> : k_step3_elem<false, false>),
                           dim3((unsigned)w.seg3_cap), dim3(256), 0, s, op, (const int *)B->rowx_ptr,
                           (const int *)B->rowx_col, (const double *)B->rowx_val, C->m, (const int4 *)w.seg3,
                           (const int *)(w.segcnt + C->tilem), C->tile_ptr, C->tile_colidx, C->mask, C->csr_ptr,
                           C->tile_nnz, C->csr_col, C->val, (const unsigned char *)w.hflag);
        TSG_LAUNCH_CHECK();
        if (w.hflag && w.nitems > 0) {
            hipLaunchKernelGGL(k_step3_heavy, dim3(w.hwaves), dim3(64), 0, s, op.Arowp, op.Acolx, op.Avalx,
                               (const int *)B->rowx_ptr, (const int *)B->rowx_col, (const double *)B->rowx_val,
                               (const int2 *)w.hitems, (const int *)w.hcount, C->tile_ptr, C->tile_colidx, C->mask,
                               C->csr_ptr, C->tile_nnz, C->val, w.cursor, w.hstride, w.hnext);
            TSG_LAUNCH_CHECK();
        }
    } else if (P.all || P.s3_unfit) {
        const unsigned nblk = tp_grid(tsg_div_up(C->numtile, S23_WAVES * TILES_PER_WAVE));
        hipLaunchKernelGGL(k_step3, dim3(nblk), dim3(64 * S23_WAVES), 0, s, op, C->tile_ptr, C->tile_rowidx,
                           C->tile_colidx, C->numtile, C->mask, C->csr_ptr, C->tile_nnz, C->csr_col, C->val,
                           P.dense_min, P.all, (const unsigned char *)w.rowcls);
        TSG_LAUNCH_CHECK();
    }
    return TSG_OK;
}

static int alloc_c_tiles(tsg_ctx *ctx, tsg_dtiles *C, long long numtileC) {
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_colidx, (size_t)numtileC + 1));
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_rowidx, (size_t)numtileC + 1));
    TSG_TRY(tsg_alloc_t(ctx, &C->mask, (size_t)numtileC * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &C->csr_ptr, (size_t)numtileC * 16 + 16));
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_nnz, (size_t)numtileC + 1));
    return TSG_OK;
}

// Temporaries of one pipeline run, freed on every exit (stream-ordered: the
// caching allocator hands a block out again only behind the work queued on
// the context stream).
struct SpgemmTemps {
    unsigned *keep = nullptr;
    void *scr = nullptr, *wbuf = nullptr, *w3buf = nullptr, *bb = nullptr;
    unsigned char *hflag = nullptr;
    int *hcnt = nullptr, *hinfo = nullptr, *hcur = nullptr, *span = nullptr;
    int2 *hitems = nullptr;
    hipEvent_t ev[6] = {};
    int nev = 0;
};

static void free_temps(tsg_ctx *ctx, SpgemmTemps &t) {
    void *ps[] = {t.keep, t.scr, t.wbuf, t.w3buf, t.hflag, t.hcnt, t.hinfo, t.hcur, t.hitems, t.span, t.bb};
    for (void *p : ps) tsg_free(ctx, p);
    for (int i = 0; i < t.nev; i++) (void)hipEventDestroy(t.ev[i]);
    t = SpgemmTemps();
}

// symbolic: steps 1 and 2 and the step-3 plan (row classes, element
// segments, heavy-row work list) only; C's values and in-tile columns are
// neither allocated nor computed.  stats then carries C's exact tile and
// nonzero counts and the heavy-row scratch the numeric step would add.
static int spgemm_run(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, const tsg_tune *tune_in,
                      tsg_dtiles *C, tsg_stats *stats, tsg_spgemm_plan *plan_out, bool symbolic,
                      SpgemmTemps &T) {
    if (!A || !B || A->n != B->m) {
        tsg_set_error("tsg_spgemm: inner dimensions differ (%d vs %d)", A ? A->n : -1, B ? B->m : -1);
        return TSG_ERR_ARG;
    }
    if (!B->csc_ptr) { tsg_set_error("tsg_spgemm: B needs its CSC tile index (tsg_tiles_build_csc)"); return TSG_ERR_STATE; }
    tsg_tune tune;
    if (tune_in) tune = *tune_in; else tsg_tune_default(&tune);
    if (tune.dense_acc_min_nnz < 1) tune.dense_acc_min_nnz = 1;
    tsg_spgemm_plan P;
    memset(&P, 0, sizeof(P));
    P.all = tune.force_tile_parallel ? 1 : 0;
    P.dense_min = tune.dense_acc_min_nnz;
    // element path: B carries its row index, and C's columns fit the
    // 28-bit entry keys of the element kernel's LDS hash
    // element-wise keys: (column offset) << 4 in 32 bits, and tile columns
    // below 2^23 - 1 for the step-2 hash slots
    P.s3_elem = B->rowx_ptr != nullptr && tune.element_mode != 1 && B->n <= (1 << 27) - 32;
    P.elem_all = B->rowx_ptr != nullptr && tune.element_mode == 2;
    hipStream_t s = ctx->stream;

    hipEvent_t *ev = T.ev;
    const bool timed = stats != nullptr;
    if (timed) {
        memset(stats, 0, sizeof(*stats));
        for (; T.nev < 6; T.nev++) TSG_HIP(hipEventCreate(&ev[T.nev]));
        TSG_HIP(hipEventRecord(ev[0], s));
        TSG_HIP(hipEventRecord(ev[4], s));
        TSG_HIP(hipEventRecord(ev[5], s));
    }
    C->m = A->m;
    C->n = B->n;
    C->tilem = A->tilem;
    C->tilen = B->tilen;
    // ---- step 1
    TSG_TRY(tsg_alloc_t(ctx, &C->tile_ptr, (size_t)C->tilem + 1));
    if (tune.s1_window_words != 0 && tune.s1_window_words != 256 && tune.s1_window_words != 1024 &&
        tune.s1_window_words != 2048) {
        tsg_set_error("spgemm: s1_window_words %d is not 0, 256, 1024 or 2048", tune.s1_window_words);
        return TSG_ERR_ARG;
    }
    P.s1w = tsg_s1_words(B, tune);
    const size_t kb = keep_bytes(C->tilem, P.s1w);
    if (kb) TSG_TRY(tsg_alloc(ctx, (void **)&T.keep, kb));
    TSG_TRY(tsg_alloc(ctx, &T.scr, 64));
    void *scr = T.scr;
    unsigned long long *bins = (unsigned long long *)scr;
    int *st2 = reinterpret_cast<int *>(static_cast<char *>(scr) + 32);  // step-2 sizing maxima
    int *bany = reinterpret_cast<int *>(static_cast<char *>(scr) + 48);  // some B tile-row bitmap kept
    TSG_HIP(hipMemsetAsync(scr, 0, 64, s));
    TSG_TRY(tsg_alloc_t(ctx, &T.span, (size_t)C->tilem + 1));
    const int4 *bdesc = nullptr;
    const unsigned *bbits = nullptr;
    if (const size_t bbb = C->tilem > 0 ? brow_bits_bytes(B) : 0) {
        TSG_TRY(tsg_alloc(ctx, &T.bb, bbb));
        TSG_TRY(launch_brow_bits(s, B, T.bb, &bdesc, &bbits, bany));
    }
    TSG_TRY(launch_step1(s, A, B, C, false, P.s1w, T.keep, bdesc, bbits, bany, T.span));
    if (C->tilem > 0) {
        hipLaunchKernelGGL(k_row_stats, dim3(std::min<unsigned>(ROW_STATS_BLOCKS, tsg_div_up(C->tilem, 1024))),
                           dim3(1024), 0, s, (const int *)C->tile_ptr, (const int *)T.span, C->tilem, st2);
        TSG_LAUNCH_CHECK();
    }
    // C's tile-row offsets; their total and the step-2 sizing maxima come
    // back in one read-back
    long long numtileC = 0;
    int smax[4] = {0, 0, 0, 0};
    {
        const long long np1 = tsg_scan_partial_count(C->tilem);
        void *part1v;
        TSG_TRY(tsg_scratch(ctx, (size_t)np1 * sizeof(long long), &part1v));
        long long *part1 = (long long *)part1v;
        TSG_TRY(tsg_scan_i32_enqueue(s, C->tile_ptr, C->tile_ptr, C->tilem, part1));
        long long *pin = ctx->pinned;
        TSG_HIP(hipMemcpyAsync(pin, part1 + np1 - 1, sizeof(long long), hipMemcpyDeviceToHost, s));
        TSG_HIP(hipMemcpyAsync(pin + 1, st2, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        TSG_HIP(hipStreamSynchronize(s));
        numtileC = pin[0];
        if (numtileC > 0x7fffffffLL) {
            tsg_set_error("prefix sum total %lld exceeds int32", numtileC);
            return TSG_ERR_OVERFLOW;
        }
        memcpy(smax, pin + 1, sizeof(smax));
    }
    tsg_free(ctx, T.span);
    T.span = nullptr;
    P.numtileC = numtileC;
    C->numtile = (int)numtileC;
    TSG_TRY(alloc_c_tiles(ctx, C, numtileC));
    size_t woff[WORK_NARR];
    const size_t wbytes = work_layout(C->tilem, numtileC, 0, woff);
    TSG_TRY(tsg_alloc(ctx, &T.wbuf, woff[WORK_NARR - 1]));  // everything but the step-3 segments
    (void)wbytes;
    void *wbuf = T.wbuf;
    SpgemmWork w = work_at(wbuf, C->tilem, numtileC, 0, nullptr);
    TSG_TRY(launch_step1(s, A, B, C, true, P.s1w, T.keep, bdesc, bbits, bany));
    tsg_free(ctx, T.keep);
    T.keep = nullptr;
    tsg_free(ctx, T.bb);
    T.bb = nullptr;
    if (timed) TSG_HIP(hipEventRecord(ev[1], s));
    // step-2 LDS sizing: the largest segment and the widest tile-row span
    P.s2u = G2_UMAX;
    P.s2map = 0;
    if (numtileC > 0) {
        const int umax = smax[0], span = smax[1], mspan = smax[2];
        P.s2u = std::max(1, umax);
        P.s2map = umax <= 256 ? mspan : 0;
        P.s3pk = span < G3E_PK_SPAN ? 1 : 0;
    }
    // ---- step 2
    const TileOperands op = make_operands(A, B);
    TSG_TRY(launch_step2(s, op, B, C, P, w));
    if (timed && numtileC > 0) {
        hipLaunchKernelGGL(k_bins, dim3(std::min<unsigned>(2048, tsg_div_up(numtileC, 256))), dim3(256), 0, s,
                           C->tile_nnz, numtileC, tune.dense_acc_min_nnz, bins);
        TSG_LAUNCH_CHECK();
    }
    // C's tile offsets, then the step-3 row classes (which need only the
    // device offsets); C's nonzero count and the class maxima come back in
    // one read-back
    long long nnzC = 0;
    const long long np2 = tsg_scan_partial_count(numtileC);
    void *part2v;
    TSG_TRY(tsg_scratch(ctx, (size_t)np2 * sizeof(long long), &part2v));
    long long *part2 = (long long *)part2v;
    TSG_TRY(tsg_scan_i32_enqueue(s, C->tile_nnz, C->tile_nnz, numtileC, part2));
    if (timed) TSG_HIP(hipEventRecord(ev[2], s));
    const bool classes = !P.all && numtileC > 0;
    if (classes) {
        TSG_TRY(launch_row_class(s, op, C, w.rowcls, w.cls_part, w.cls_out, P.elem_all));
    }
    {
        long long *pin = ctx->pinned;  // [0] the total, [1..4] the eight class words
        TSG_HIP(hipMemcpyAsync(pin, part2 + np2 - 1, sizeof(long long), hipMemcpyDeviceToHost, s));
        if (classes) TSG_HIP(hipMemcpyAsync(pin + 1, w.cls_out, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
        TSG_HIP(hipStreamSynchronize(s));
        nnzC = pin[0];
        if (nnzC > 0x7fffffffLL) {
            tsg_set_error("prefix sum total %lld exceeds int32", nnzC);
            return TSG_ERR_OVERFLOW;
        }
        if (classes) {
            const int *h = reinterpret_cast<const int *>(pin + 1);
            P.s3_unfit = h[0] > 0;
            P.s3_rows = h[1] > 0;
            for (int i = 0; i < 6; i++) P.g3[i] = std::max(1, h[1 + i]);
        }
    }
    P.nnzC = nnzC;
    C->nnz = (int)nnzC;
    // ---- step 3
    if (!symbolic) {
        TSG_TRY(tsg_alloc_t(ctx, &C->csr_col, (size_t)nnzC + 1));
        TSG_TRY(tsg_alloc_t(ctx, &C->val, (size_t)nnzC + 1));
    }
    if (nnzC > 0) {
        if (!P.all && P.s3_unfit && P.s3_elem) {
            const SpgemmWork w3 = work_at(wbuf, C->tilem, numtileC, nnzC, nullptr);
            TSG_TRY(tsg_alloc(ctx, &T.w3buf, 16 * (size_t)w3.seg3_cap));
            w.seg3 = (int4 *)T.w3buf;
            w.seg3_cap = w3.seg3_cap;
            if (op.Arowp && !tune.no_heavy_rows) {
                // rows with very many products: one wave each (k_step3_heavy)
                const int m = C->m;
                TSG_TRY(tsg_alloc_t(ctx, &T.hflag, (size_t)m + 16));
                TSG_TRY(tsg_alloc_t(ctx, &T.hcnt, (size_t)m + 1));
                TSG_TRY(tsg_alloc_t(ctx, &T.hinfo, 4));  // {longest heavy A row, heavy rows, item counter, pad}
                unsigned char *hflag = T.hflag;
                int *hcnt = T.hcnt, *hinfo = T.hinfo;
                TSG_HIP(hipMemsetAsync(hinfo, 0, 16, s));
                hipLaunchKernelGGL(k_heavy_flags, dim3(tsg_div_up(m, 256)), dim3(256), 0, s, op.Arowp, op.Acolx,
                                   (const int *)B->rowx_ptr, (const unsigned char *)w.rowcls, C->tile_ptr,
                                   C->tile_colidx, m,
                                   (long long)(tune.heavy_min_products > 0 ? tune.heavy_min_products
                                                                           : HEAVY_MIN_PRODUCTS),
                                   (long long)(tune.heavy_item_products > 0 ? tune.heavy_item_products
                                                                            : HEAVY_ITEM_PRODUCTS),
                                   hflag, hcnt, hinfo);
                TSG_LAUNCH_CHECK();
                long long ni = 0;
                TSG_TRY(tsg_scan_i32(ctx, hcnt, hcnt, m, &ni));
                if (ni > 0x7fffffffLL) {
                    tsg_set_error("spgemm: %lld heavy-row work items exceed the int32 range", ni);
                    return TSG_ERR_ARG;
                }
                if (ni > 0) {
                    long long info = 0;  // {longest heavy A row, heavy rows}
                    TSG_TRY(tsg_read_scalar_i64(ctx, hinfo, 1, &info));
                    const long long maxdeg = info & 0xffffffffLL, nrows = info >> 32;
                    TSG_TRY(tsg_alloc_t(ctx, &T.hitems, (size_t)ni));
                    int2 *hitems = T.hitems;
                    hipLaunchKernelGGL(k_heavy_list, dim3(tsg_div_up(m, 256)), dim3(256), 0, s, (const int *)hcnt,
                                       (const unsigned char *)hflag, C->tile_ptr, C->tile_colidx, m, hitems);
                    TSG_LAUNCH_CHECK();
                    // one wave per item up to HEAVY_MAX_WAVES, cursors within HEAVY_CURSOR_BYTES
                    const long long by_mem = std::max<long long>(256, HEAVY_CURSOR_BYTES / (4 * std::max(1LL, maxdeg)));
                    w.hwaves = (int)std::min<long long>(ni, std::min<long long>(HEAVY_MAX_WAVES, by_mem));
                    w.hstride = std::max(1LL, maxdeg);
                    if (!symbolic) TSG_TRY(tsg_alloc_t(ctx, &T.hcur, (size_t)w.hwaves * (size_t)w.hstride));
                    int *hcur = T.hcur;
                    w.hflag = hflag;
                    w.hitems = hitems;
                    w.hcount = hcnt + m;
                    w.nheavy = (int)nrows;
                    w.nitems = (int)ni;
                    w.cursor = hcur;
                    w.hnext = hinfo + 2;
                }
            }
        }
        if (!symbolic) TSG_TRY(launch_step3(s, op, B, C, P, w, timed ? ev + 4 : nullptr));
    }
    if (timed) {
        TSG_HIP(hipEventRecord(ev[3], s));
        TSG_HIP(hipEventSynchronize(ev[3]));
        TSG_TRY(ev_ms(ev[0], ev[1], &stats->t_step1_ms));
        TSG_TRY(ev_ms(ev[1], ev[2], &stats->t_step2_ms));
        TSG_TRY(ev_ms(ev[2], ev[3], &stats->t_step3_ms));
        TSG_TRY(ev_ms(ev[0], ev[3], &stats->t_total_ms));
        TSG_TRY(ev_ms(ev[4], ev[5], &stats->t_step3_rows_ms));
        unsigned long long hb[4];
        TSG_HIP(hipMemcpyAsync(hb, bins, sizeof(hb), hipMemcpyDeviceToHost, s));
        TSG_HIP(hipStreamSynchronize(s));
        stats->tiles_empty = (long long)hb[0];
        stats->tiles_sparse_acc = (long long)hb[1];
        stats->tiles_dense_acc = (long long)hb[2];
        stats->numtileC = numtileC;
        stats->nnzC = nnzC;
        stats->heavy_rows = w.nheavy;
        stats->heavy_bytes = (long long)w.hwaves * w.hstride * 4 + (w.nheavy ? 5LL * C->m + 8LL * w.nitems : 0);
    }
    if (plan_out) *plan_out = P;
    return TSG_OK;
}

int tsg_spgemm_planned(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, const tsg_tune *tune_in,
                       tsg_dtiles *C, tsg_stats *stats, tsg_spgemm_plan *plan_out, bool symbolic) {
    memset(C, 0, sizeof(*C));
    SpgemmTemps T;
    const int st = spgemm_run(ctx, A, B, tune_in, C, stats, plan_out, symbolic, T);
    free_temps(ctx, T);
    if (st != TSG_OK) tsg_tiles_free(ctx, C);  // a partly built C goes back too
    return st;
}


extern "C" int tsg_spgemm(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, const tsg_tune *tune_in,
                          tsg_dtiles *C, tsg_stats *stats) {
    return tsg_spgemm_planned(ctx, A, B, tune_in, C, stats, nullptr);
}

size_t tsg_spgemm_work_bytes(const tsg_dtiles *C, const tsg_dtiles *B, int s1w) {
    size_t off[WORK_NARR];
    return work_layout(C->tilem, C->numtile, C->nnz, off, s1w, brow_bits_bytes(B));
}

int tsg_spgemm_enqueue(hipStream_t s, const tsg_dtiles *A, const tsg_dtiles *B, tsg_dtiles *C,
                       const tsg_spgemm_plan &P, void *work) {
    const SpgemmWork w = work_at(work, C->tilem, C->numtile, C->nnz, nullptr, P.s1w, brow_bits_bytes(B));
    const int4 *bdesc = nullptr;
    const unsigned *bbits = nullptr;
    if (C->tilem > 0) TSG_TRY(launch_brow_bits(s, B, w.bbits, &bdesc, &bbits, nullptr));
    TSG_TRY(launch_step1(s, A, B, C, false, P.s1w, w.bmstore, bdesc, bbits, nullptr));
    TSG_TRY(tsg_scan_i32_enqueue(s, C->tile_ptr, C->tile_ptr, C->tilem, w.part1));
    TSG_TRY(launch_step1(s, A, B, C, true, P.s1w, w.bmstore, bdesc, bbits, nullptr));
    const TileOperands op = make_operands(A, B);
    TSG_TRY(launch_step2(s, op, B, C, P, w));
    TSG_TRY(tsg_scan_i32_enqueue(s, C->tile_nnz, C->tile_nnz, C->numtile, w.part2));
    if (C->nnz > 0) {
        if (!P.all) TSG_TRY(launch_row_class(s, op, C, w.rowcls, w.cls_part, w.cls_out, P.elem_all));
        TSG_TRY(launch_step3(s, op, B, C, P, w));
    }
    return TSG_OK;
}

// ------------------------------------------------------------------ chunks
// Row lengths of B from its tile masks: row 16K+r is the popcount of row r
// over the tiles of tile-row K.
__global__ void __launch_bounds__(256)
k_tile_row_lengths(const int *__restrict__ Bptr, const unsigned short *__restrict__ Bmask, int m,
                   int *__restrict__ len) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int K = i >> 4, r = i & 15;
    int c = 0;
    for (int t = Bptr[K]; t < Bptr[K + 1]; t++) c += __popc((unsigned)Bmask[(long long)t * 16 + r]);
    len[i] = c;
}

// Intermediate products of every A tile-row: sum over A's nonzeros (i, k)
// of B's row length k.  One wave per tile-row, lanes over its tiles.
__global__ void __launch_bounds__(256)
k_tile_row_products(const int *__restrict__ Aptr, const int *__restrict__ Acol, const unsigned short *__restrict__ Amask,
                    int tilem, const int *__restrict__ blen, int bm, unsigned long long *__restrict__ out) {
    const int I = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (I >= tilem) return;
    unsigned long long s = 0;
    for (int t = Aptr[I] + lane; t < Aptr[I + 1]; t += 64) {
        const int k0 = Acol[t] * 16;
        const uint4 *m4 = reinterpret_cast<const uint4 *>(Amask + (long long)t * 16);
        const uint4 a = m4[0], b = m4[1];
        const unsigned w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        int cc[16];
#pragma unroll
        for (int c = 0; c < 16; c++) cc[c] = 0;
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int c = 0; c < 16; c++) cc[c] += ((w[q] >> (15 - c)) & 1u) + ((w[q] >> (31 - c)) & 1u);
#pragma unroll
        for (int c = 0; c < 16; c++)
            if (cc[c] && k0 + c < bm) s += (unsigned long long)cc[c] * (unsigned long long)blen[k0 + c];
    }
    s = tsg_wave_sum(s);
    if (lane == 0) out[I] = s;
}

int tsg_tile_row_products(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, unsigned long long *d_out) {
    if (A->n != B->m) { tsg_set_error("tile_row_products: inner dimensions differ"); return TSG_ERR_ARG; }
    int *blen;
    TSG_TRY(tsg_alloc_t(ctx, &blen, (size_t)B->m + 1));
    if (B->m > 0) {
        hipLaunchKernelGGL(k_tile_row_lengths, dim3(tsg_div_up(B->m, 256)), dim3(256), 0, ctx->stream, B->tile_ptr,
                           B->mask, B->m, blen);
        TSG_LAUNCH_CHECK();
    }
    if (A->tilem > 0) {
        hipLaunchKernelGGL(k_tile_row_products, dim3(tsg_div_up(A->tilem, 4)), dim3(256), 0, ctx->stream, A->tile_ptr,
                           A->tile_colidx, A->mask, A->tilem, blen, B->m, d_out);
        TSG_LAUNCH_CHECK();
    }
    tsg_free(ctx, blen);
    return TSG_OK;
}

// Tile-row ranges of A whose intermediate products stay within
// max_products each (a single tile-row above it forms its own chunk).
// Every chunk's C then has at most max_products nonzeros.
extern "C" int tsg_spgemm_chunk_plan(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, long long max_products,
                                     int max_chunks, int *bounds, int *nchunks, unsigned long long *products_total) {
    if (!A || !B || !bounds || !nchunks || max_chunks < 1 || max_products < 1) {
        tsg_set_error("tsg_spgemm_chunk_plan: bad arguments");
        return TSG_ERR_ARG;
    }
    unsigned long long *d;
    TSG_TRY(tsg_alloc_t(ctx, &d, (size_t)A->tilem + 1));
    TSG_TRY(tsg_tile_row_products(ctx, A, B, d));
    std::vector<unsigned long long> h((size_t)A->tilem + 1);
    if (A->tilem > 0)
        TSG_HIP(hipMemcpyAsync(h.data(), d, 8 * (size_t)A->tilem, hipMemcpyDeviceToHost, ctx->stream));
    TSG_HIP(hipStreamSynchronize(ctx->stream));
    tsg_free(ctx, d);
    int n = 0;
    bounds[0] = 0;
    unsigned long long acc = 0, tot = 0;
    for (int I = 0; I < A->tilem; I++) {
        tot += h[I];
        if (acc > 0 && acc + h[I] > (unsigned long long)max_products) {
            if (n + 1 >= max_chunks) {
                tsg_set_error("tsg_spgemm_chunk_plan: more than %d chunks needed", max_chunks);
                return TSG_ERR_ARG;
            }
            bounds[++n] = I;
            acc = 0;
        }
        acc += h[I];
    }
    bounds[++n] = A->tilem;
    *nchunks = n;
    if (products_total) *products_total = tot;
    return TSG_OK;
}

// C rows [16*r0, min(16*r1, m)) of A*B: A restricted to tile-rows [r0, r1)
// as a view (tile_ptr offset into A's, every other array shared).
extern "C" int tsg_spgemm_tile_rows(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, int r0, int r1,
                                    const tsg_tune *tune, tsg_dtiles *C, tsg_stats *stats) {
    return tsg_spgemm_tile_rows_ex(ctx, A, B, r0, r1, tune, C, stats, false);
}

int tsg_spgemm_tile_rows_ex(tsg_ctx *ctx, const tsg_dtiles *A, const tsg_dtiles *B, int r0, int r1,
                            const tsg_tune *tune, tsg_dtiles *C, tsg_stats *stats, bool symbolic) {
    if (!A || r0 < 0 || r1 < r0 || r1 > A->tilem) {
        memset(C, 0, sizeof(*C));
        tsg_set_error("tsg_spgemm_tile_rows: bad tile-row range [%d, %d)", r0, r1);
        return TSG_ERR_ARG;
    }
    tsg_dtiles v = *A;
    v.tile_ptr = A->tile_ptr + r0;
    if (A->rowx_ptr) v.rowx_ptr = A->rowx_ptr + 16 * (long long)r0;  // indices stay absolute
    v.tilem = r1 - r0;
    v.m = std::min(16 * r1, A->m) - 16 * r0;
    if (v.m < 0) v.m = 0;
    return tsg_spgemm_planned(ctx, &v, B, tune, C, stats, nullptr, symbolic);
}
