// nb_keys.hip -- stage 2, the octree build (buildtree.cl:12-200), from sorted
// octant-path keys instead of insertion.
//
// A body's place in the reference's octree is fixed by its octant path: the
// octant at each level, taken against cell centres that the reference derives
// from the root with one FP32 recurrence (buildtree.cl:124-136), and a body
// sits in the first level where no other body shares its path. So every body
// works out its own path with that same recurrence, 3 bits per level, and the
// bodies are radix-sorted by it. The cells are the path prefixes that at
// least two bodies share; from the sorted keys they follow without any
// pointer chasing:
//   * a binary radix tree over the sorted keys (Karras, "Maximizing
//     parallelism in the construction of BVHs, octrees, and k-d trees", HPG
//     2012): internal node i finds its key range and split by binary searches
//     on delta(i, j), the common prefix length of keys i and j in bits;
//   * an octree cell of level k is the prefix of 3k bits: binary node X holds
//     the cells of the levels k with 3k in (delta(parent(X)), delta(X)] -- a
//     chain of them when the bodies of X share several more octants, exactly
//     as the reference's split loop creates them (buildtree.cl:106-165);
//   * the parent of X's top cell is the deepest cell of the nearest binary
//     ancestor that holds any, and it sits in the slot of its octant at the
//     parent's level; a body sits in the deepest cell of its nearest
//     ancestor that holds one.
// The result is written in the layout the rest of the pipeline reads:
// epoch-tagged child slots (a slot of an older epoch reads as empty, so the
// last step's tree needs no clearing) and per-cell records, cells numbered in
// key order with the root as cell 0. The sorted body order is the
// octant-path (Morton) order that the sort stage produces (sort.cl:12-76).
//
// Keys hold 21 levels in 63 bits (level 0 in the top bits). Bodies that share
// all 21 (rare: 2 of 4M in the clustered benchmark universe) form runs of
// equal keys; one thread per run extends its bodies' paths to 84 levels and
// orders the run by them (in registers up to kMaxRun bodies, through global
// memory beyond). Limit: cells on levels below kBuildLevelLimit (bodies at
// most that deep); beyond it the build stops with a depth error, and a tree
// of more cells than the pool holds stops it with a pool error.
#include "nb_device.h"

namespace nb {

constexpr int kKeyLevels = 21; // levels per 64-bit key word (63 bits)
constexpr int kDeepWords = 3;  // key words past the first: levels 21 .. 83
constexpr int kMaxRun = 16;    // bodies sharing the first 21 levels, per run

// a child slot of the current tree: (epoch << 32) | child (nb_tree.hip reads them)
__device__ __forceinline__ unsigned long long slot_tag(uint32_t ep, int v)
{
    return ((unsigned long long)ep << 32) | (uint32_t)v;
}

__device__ __forceinline__ int key_octant(float cx, float cy, float cz, float4 p)
{
    return (cx < p.x ? 1 : 0) | (cy < p.y ? 2 : 0) | (cz < p.z ? 4 : 0);
}

// the child centre of octant `o` (buildtree.cl:124-136: x = parent - r/2 +
// (bit ? r : 0), r the parent's half size)
__device__ __forceinline__ void key_descend(float &cx, float &cy, float &cz, float &r, int o)
{
    const float rh = r * 0.5f;
    cx = (cx - rh) + ((o & 1) ? r : 0.0f);
    cy = (cy - rh) + ((o & 2) ? r : 0.0f);
    cz = (cz - rh) + ((o & 4) ? r : 0.0f);
    r = rh;
}

// ---------------------------------------------------------------------------
// Keys: 21 levels of each body's octant path, level 0 in bits 62..60, for the
// bodies in the order `order` (the identity when null).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_keys(Dev d, const int *__restrict__ order, unsigned long long *key, int *val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const Ctl *ctl = d.ctl;
    if (i >= d.n || ctl->error) return;
    const int b = order ? order[i] : i;
    const float4 p = d.pm4[b];
    float cx = ctl->root[0], cy = ctl->root[1], cz = ctl->root[2], r = ctl->radius;
    unsigned long long k = 0;
#pragma unroll
    for (int L = 0; L < kKeyLevels; ++L) {
        const int o = key_octant(cx, cy, cz, p);
        k = (k << 3) | (unsigned long long)o;
        key_descend(cx, cy, cz, r, o);
    }
    key[i] = k;
    val[i] = b;
}

// ---------------------------------------------------------------------------
// LSD radix sort of (key, body) by 8-bit digits, 8 passes. A tile of
// kRsTile keys per block: per-tile digit counts (k_rs_count), their
// digit-major exclusive offsets over the tiles (k_rs_offsets), then each tile
// ranks its keys stably, sorts itself by digit in LDS and writes each digit's
// run out contiguously (k_rs_scatter).
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsPer = 8;
constexpr int kRsTile = kRsThreads * kRsPer; // 2048 keys
constexpr int kRsBins = 256;
static_assert(kRsTile == kKeyTile, "the engine sizes the per-tile arrays with kKeyTile");

__global__ __launch_bounds__(kRsThreads) void k_rs_count(const unsigned long long *__restrict__ key, int n, int shift,
                                                         unsigned *__restrict__ hist, int ntiles)
{
    __shared__ unsigned h[kRsBins];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
#pragma unroll
    for (int j = 0; j < kRsPer; ++j) {
        const int idx = base + j * kRsThreads + t;
        if (idx < n) atomicAdd(&h[(unsigned)(key[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)t * ntiles + blockIdx.x] = h[t];
}

// Block b: the exclusive prefix of digit b's counts over the tiles (into
// offs) and the digit's total. 1024 threads, each a contiguous run of tiles.
__global__ __launch_bounds__(1024) void k_rs_offsets(const unsigned *__restrict__ hist, int ntiles,
                                                     unsigned *__restrict__ offs, unsigned *__restrict__ total)
{
    __shared__ unsigned s[1024];
    const int t = threadIdx.x, b = blockIdx.x;
    const int per = (ntiles + 1023) / 1024;
    const unsigned *h = hist + (size_t)b * ntiles;
    unsigned *o = offs + (size_t)b * ntiles;
    unsigned sum = 0;
    for (int k = 0; k < per; ++k) {
        const int tile = t * per + k;
        if (tile < ntiles) sum += h[tile];
    }
    s[t] = sum;
    __syncthreads();
    for (int st = 1; st < 1024; st <<= 1) {
        const unsigned u = t >= st ? s[t - st] : 0u;
        __syncthreads();
        s[t] += u;
        __syncthreads();
    }
    unsigned run = s[t] - sum; // exclusive
    for (int k = 0; k < per; ++k) {
        const int tile = t * per + k;
        if (tile < ntiles) {
            o[tile] = run;
            run += h[tile];
        }
    }
    if (t == 1023) total[b] = s[1023];
}

__device__ __forceinline__ unsigned long long lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// exclusive scan of one value per thread over a 256-thread block (v[t] in,
// the result returned); `s` is 256 words of LDS
__device__ __forceinline__ unsigned block_excl_scan256(unsigned v, unsigned *s)
{
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int st = 1; st < 256; st <<= 1) {
        const unsigned u = t >= st ? s[t - st] : 0u;
        __syncthreads();
        s[t] += u;
        __syncthreads();
    }
    const unsigned r = s[t] - v;
    __syncthreads();
    return r;
}

// One tile. A key's place in the tile sorted by digit = the tile's keys of
// smaller digits + its rank among the keys of its digit, counted in input
// order: keys of its digit in earlier rounds of 256, in earlier waves of this
// round, and in lower lanes of its wave (from eight ballots of the digit's
// bits). The tile is sorted in LDS, then written out in tile order, so that
// consecutive threads write consecutive places of each digit's run.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const unsigned long long *__restrict__ kin,
                                                           const int *__restrict__ vin,
                                                           unsigned long long *__restrict__ kout,
                                                           int *__restrict__ vout, int n, int shift,
                                                           const unsigned *__restrict__ hist,
                                                           const unsigned *__restrict__ offs,
                                                           const unsigned *__restrict__ total, int ntiles)
{
    __shared__ unsigned long long sk[kRsTile];
    __shared__ int sv[kRsTile];
    __shared__ unsigned gstart[kRsBins]; // global place of this tile's first key of each digit
    __shared__ unsigned lstart[kRsBins]; // its place in the tile
    __shared__ unsigned run[kRsBins];    // keys of each digit placed so far
    __shared__ unsigned cnt[4][kRsBins]; // per wave and digit, this round
    __shared__ unsigned scratch[kRsBins];
    const int t = threadIdx.x, w = t >> 6;
    const int tile = blockIdx.x, base = tile * kRsTile;
    const int tcount = min(kRsTile, n - base);
    const unsigned tot = total[t];
    const unsigned digit_base = block_excl_scan256(tot, scratch);
    gstart[t] = digit_base + offs[(size_t)t * ntiles + tile];
    lstart[t] = block_excl_scan256(hist[(size_t)t * ntiles + tile], scratch);
    run[t] = 0;
    cnt[0][t] = cnt[1][t] = cnt[2][t] = cnt[3][t] = 0;
    __syncthreads();
    const unsigned long long below = lanemask_lt();
    for (int j = 0; j < kRsPer; ++j) {
        const int li = j * kRsThreads + t;
        const bool valid = li < tcount;
        const unsigned long long k = valid ? kin[base + li] : 0ull;
        const int v = valid ? vin[base + li] : 0;
        const unsigned dg = (unsigned)(k >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bal = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bal : ~bal;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) cnt[w][dg] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned at = lstart[dg] + run[dg] + (unsigned)rank;
            for (int ww = 0; ww < w; ++ww) at += cnt[ww][dg];
            sk[at] = k;
            sv[at] = v;
        }
        __syncthreads();
        {
            run[t] += cnt[0][t] + cnt[1][t] + cnt[2][t] + cnt[3][t];
            cnt[0][t] = cnt[1][t] = cnt[2][t] = cnt[3][t] = 0;
        }
        __syncthreads();
    }
    for (int li = t; li < tcount; li += kRsThreads) {
        const unsigned long long k = sk[li];
        const unsigned dg = (unsigned)(k >> shift) & 255u;
        const unsigned dest = gstart[dg] + ((unsigned)li - lstart[dg]);
        kout[dest] = k;
        vout[dest] = sv[li];
    }
}

// ---------------------------------------------------------------------------
// Warm sort. After the first step the bodies come in the last tree's order
// (sorted[]), which the new keys nearly keep, and the last tree's sorted keys
// are still in kb_key[0]. Bucketing by splitters taken from those keys then
// moves most bodies a short way and every bucket is small:
//   * splitters: every kBucketT-th of the last tree's sorted keys;
//   * a body's bucket: the number of splitters <= its key. A block's bodies
//     are consecutive in the last order, so their buckets are nearly all in
//     a window of kWin around the one their places suggest: searched in an
//     LDS copy of the window's splitters, counted in LDS, one atomic per
//     bucket and block (k_keys_warm, with the keys); the counts' exclusive
//     scan (k_bucket_scan); then each body takes the next place of its
//     bucket the same way (k_bucket_scatter);
//   * each bucket is sorted by one wave (k_bucket_sort): up to kBucketCap
//     bodies held in registers, stable radix passes over the key bits that
//     vary inside it, permuted through LDS; a larger bucket (stale order) is
//     listed and sorted by a block through device memory (k_bucket_sort_big).
// The order inside a bucket before its sort depends on the atomics; the keys
// are distinct after the run ordering below (equal keys are ordered by their
// deeper levels), so the result does not.
// ---------------------------------------------------------------------------
constexpr int kBsThreads = 512; // blocks of the large-bucket sort
constexpr int kBsWaves = kBsThreads / 64;
static_assert(kBucketCap % 64 == 0, "whole rows of a wave");

// Splitters spl[0 .. S-1]: the keys at positions j * kBucketT of `key` (the
// last tree's sorted keys, so non-decreasing: a decrease means the buffer does
// not hold them and stops the build with an error); the bucket counters and
// the list of large buckets cleared.
__global__ __launch_bounds__(256) void k_splitters(Ctl *ctl, const unsigned long long *__restrict__ key, int S,
                                                   unsigned long long *__restrict__ spl, unsigned *__restrict__ bcnt,
                                                   unsigned *__restrict__ bfill, int *__restrict__ big)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl->error || j > S) return;
    if (j < S) {
        const unsigned long long k = key[(size_t)j * kBucketT];
        spl[j] = k;
        if (j > 0 && key[(size_t)(j - 1) * kBucketT] > k) flag_error(ctl, kErrSync);
    }
    bcnt[j] = 0;
    bfill[j] = 0;
    if (j == 0) big[0] = 0;
}

// the first j with spl[j] > k (S if none): spl is non-decreasing; searched
// outward from g, then by bisection (global memory). Invariant: (lo < 0 or spl[lo] <= k) and
// (hi >= S or spl[hi] > k).
__device__ __forceinline__ int bucket_of(const unsigned long long *__restrict__ spl, int S, unsigned long long k, int g)
{
    g = min(max(g, 0), S);
    int lo, hi;
    if (g < S && spl[g] <= k) {
        lo = g;
        hi = g + 1;
        for (int step = 1; hi < S && spl[hi] <= k; hi = lo + step) {
            lo = hi;
            step <<= 1;
        }
        hi = min(hi, S);
    } else {
        hi = g;
        lo = g - 1;
        for (int step = 1; lo >= 0 && spl[lo] > k; lo = hi - step) {
            hi = lo;
            step <<= 1;
        }
        lo = max(lo, -1);
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (spl[mid] <= k) lo = mid;
        else hi = mid;
    }
    return hi;
}

// Bucket windows: a block's bodies (consecutive in the last tree's order)
// expect buckets near first / kBucketT + 1; buckets w0 .. w0 + kWin - 1 (w0
// = that - kWin / 4) are searched in an LDS copy of their splitters and
// counted in LDS bins; a body outside the window is searched in global
// memory and goes to its bucket's counter directly.
constexpr int kWinPer = 8;
constexpr int kWinBodies = 256 * kWinPer;
constexpr int kWin = 64;

// ws[0 .. kWin]: spl[w0 - 1 .. w0 + kWin - 1], 0 below the first splitter
// and ~0 past the last (keys use 63 bits); block-wide, ends with a barrier
__device__ __forceinline__ void load_window(const unsigned long long *__restrict__ spl, int S, int w0,
                                            unsigned long long *ws)
{
    const int t = threadIdx.x;
    if (t <= kWin) {
        const int j = w0 - 1 + t;
        ws[t] = j < 0 ? 0ull : (j >= S ? ~0ull : spl[j]);
    }
    __syncthreads();
}

// bucket of k if it is one of w0 .. w0 + kWin - 1, else -1
__device__ __forceinline__ int bucket_in_window(const unsigned long long *ws, int w0, unsigned long long k)
{
    if (k < ws[0] || k >= ws[kWin]) return -1;
    int lo = 0, hi = kWin; // ws[lo] <= k < ws[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ws[mid] <= k) lo = mid;
        else hi = mid;
    }
    return w0 - 1 + hi;
}

// The place of this lane among the lanes of its wave that take LDS bin o
// (those with `take`), after h[o], which moves on by their count: one LDS
// atomic per bin present in the wave instead of one per lane (the lanes of a
// wave nearly all share a bin). Every lane of the wave calls it.
__device__ __forceinline__ unsigned claim_bin(unsigned *h, int o, bool take)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned long long todo = __ballot(take);
    unsigned mine = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int bb = __shfl(o, leader);
        const unsigned long long same = __ballot(o == bb) & todo;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&h[bb], (unsigned)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) mine = base + (unsigned)__popcll(same & below);
        todo &= ~same;
    }
    return mine;
}

// Warm keys: body order[i] (the last tree's order) at place i: its key and
// value, its bucket (bid) and the bucket counts. Its record comes from
// `staged` (the records in that order, as the last fused update wrote them)
// when given, else from pm4. A block takes kWarmPer * 256 places, whose
// buckets are expected near first / kBucketT + 1; the window starts kWin / 4
// below. Every load is issued before the arithmetic.
constexpr int kWarmPer = 4;
constexpr int kWarmBodies = 256 * kWarmPer;

__global__ __launch_bounds__(256) void k_keys_warm(Dev d, const int *__restrict__ order,
                                                   const float4 *__restrict__ staged,
                                                   const unsigned long long *__restrict__ spl, int S,
                                                   unsigned long long *__restrict__ key, int *__restrict__ val,
                                                   int *__restrict__ bid, unsigned *__restrict__ bcnt)
{
    __shared__ unsigned h[kWin];
    __shared__ unsigned long long ws[kWin + 1];
    const int t = threadIdx.x, first = blockIdx.x * kWarmBodies;
    const int n = d.n;
    const Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const int w0 = first / kBucketT + 1 - kWin / 4;
    if (t < kWin) h[t] = 0;
    int b[kWarmPer];
    float4 p[kWarmPer];
#pragma unroll
    for (int j = 0; j < kWarmPer; ++j) {
        const int ic = min(first + j * 256 + t, n - 1);
        b[j] = order[ic];
        p[j] = staged ? staged[ic] : d.pm4[b[j]];
    }
    load_window(spl, S, w0, ws);
    const float rx = ctl->root[0], ry = ctl->root[1], rz = ctl->root[2], rr = ctl->radius;
#pragma unroll
    for (int j = 0; j < kWarmPer; ++j) {
        const int i = first + j * 256 + t;
        const bool valid = i < n;
        unsigned long long k = 0;
        float cx = rx, cy = ry, cz = rz, r = rr;
#pragma unroll
        for (int L = 0; L < kKeyLevels; ++L) {
            const int o = key_octant(cx, cy, cz, p[j]);
            k = (k << 3) | (unsigned long long)o;
            key_descend(cx, cy, cz, r, o);
        }
        int bk = valid ? bucket_in_window(ws, w0, k) : -1;
        const bool inwin = bk >= 0;
        if (valid && !inwin) {
            bk = bucket_of(spl, S, k, i / kBucketT + 1);
            atomicAdd(&bcnt[bk], 1u);
        }
        (void)claim_bin(h, inwin ? bk - w0 : 0, inwin);
        if (valid) {
            key[i] = k;
            val[i] = b[j];
            bid[i] = bk;
        }
    }
    __syncthreads();
    if (t < kWin && h[t]) atomicAdd(&bcnt[w0 + t], h[t]);
}

// exclusive scan of the S + 1 bucket counts: bstart[0 .. S+1] (one block).
// Chunks of kBsChunk counts go through LDS (coalesced loads and stores; the
// 8 of a thread padded to 9 words, so its reads hit distinct banks), each
// thread scans 8 consecutive counts, and the chunk's carry moves on.
constexpr int kBsPer = 8;
constexpr int kBsChunk = 1024 * kBsPer;

__device__ __forceinline__ int bs_pad(int i) { return i + i / kBsPer; }

__global__ __launch_bounds__(1024) void k_bucket_scan(const Ctl *ctl, const unsigned *__restrict__ bcnt, int S,
                                                      unsigned *__restrict__ bstart)
{
    __shared__ unsigned c[kBsChunk + kBsChunk / kBsPer];
    __shared__ unsigned wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (ctl->error) return;
    const int nb = S + 1;
    unsigned carry = 0;
    for (int c0 = 0; c0 < nb; c0 += kBsChunk) {
#pragma unroll
        for (int k = 0; k < kBsPer; ++k) {
            const int i = k * 1024 + t;
            c[bs_pad(i)] = c0 + i < nb ? bcnt[c0 + i] : 0u;
        }
        __syncthreads();
        unsigned v[kBsPer], sum = 0;
#pragma unroll
        for (int k = 0; k < kBsPer; ++k) {
            v[k] = c[bs_pad(t * kBsPer + k)];
            sum += v[k];
        }
        unsigned x = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        unsigned below = 0, total = 0;
#pragma unroll
        for (int ww = 0; ww < 16; ++ww) {
            const unsigned ws = wsum[ww];
            below += ww < w ? ws : 0u;
            total += ws;
        }
        unsigned run = carry + below + x - sum;
#pragma unroll
        for (int k = 0; k < kBsPer; ++k) {
            c[bs_pad(t * kBsPer + k)] = run;
            run += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kBsPer; ++k) {
            const int i = k * 1024 + t;
            if (c0 + i < nb) bstart[c0 + i] = c[bs_pad(i)];
        }
        carry += total;
        __syncthreads();
    }
    if (t == 0) bstart[nb] = carry;
}

// Each body takes the next place of its bucket: LDS counters for the buckets
// of the block's window (one global atomic per bucket and block), a global
// atomic for a bucket outside it.
__global__ __launch_bounds__(256) void k_bucket_scatter(const Ctl *ctl, const unsigned long long *__restrict__ kin,
                                                        const int *__restrict__ vin, const int *__restrict__ bid,
                                                        unsigned long long *__restrict__ kout, int *__restrict__ vout,
                                                        int n, const unsigned *__restrict__ bstart,
                                                        unsigned *__restrict__ bfill)
{
    __shared__ unsigned h[kWin], hb[kWin];
    if (ctl->error) return;
    const int t = threadIdx.x, first = blockIdx.x * kWinBodies;
    const int w0 = first / kBucketT + 1 - kWin / 4;
    if (t < kWin) h[t] = 0;
    unsigned long long k[kWinPer];
    int v[kWinPer], b[kWinPer];
#pragma unroll
    for (int j = 0; j < kWinPer; ++j) {
        const int i = min(first + j * 256 + t, n - 1);
        k[j] = kin[i];
        v[j] = vin[i];
        b[j] = bid[i];
    }
    __syncthreads();
    unsigned at[kWinPer];
#pragma unroll
    for (int j = 0; j < kWinPer; ++j) {
        const bool valid = first + j * 256 + t < n;
        const unsigned o = (unsigned)(b[j] - w0);
        const bool inwin = valid && o < (unsigned)kWin;
        at[j] = claim_bin(h, inwin ? (int)o : 0, inwin);
        if (valid && !inwin) at[j] = bstart[b[j]] + atomicAdd(&bfill[b[j]], 1u);
    }
    __syncthreads();
    if (t < kWin) hb[t] = h[t] ? bstart[w0 + t] + atomicAdd(&bfill[w0 + t], h[t]) : 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kWinPer; ++j) {
        if (first + j * 256 + t >= n) continue;
        const unsigned o = (unsigned)(b[j] - w0);
        const unsigned pos = o < (unsigned)kWin ? hb[o] + at[j] : at[j];
        kout[pos] = k[j];
        vout[pos] = v[j];
    }
}

struct RankLds {
    unsigned cnt[kBsWaves][kRsBins]; // per wave and digit in a round; then their exclusive prefix over the waves
    unsigned lstart[kRsBins];        // first place of each digit
    unsigned run[kRsBins];           // places of each digit taken by earlier rounds
    unsigned tot[kRsBins];           // digit counts; then each round's per-digit total
    unsigned wsum[kBsWaves];
};

// exclusive scan over the first 256 threads of the block (each passes v; the
// others pass anything and ignore the result); every thread must call it
__device__ __forceinline__ unsigned scan256_excl(unsigned v, unsigned *wsum)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned add = 0;
    for (int ww = 0; ww < w && ww < 4; ++ww) add += wsum[ww];
    __syncthreads();
    return add + x - v;
}

// One stable pass over m keys by the 8 bits at `shift`: (sk, sv) -> (dk, dv),
// all m places of dk/dv written. Rounds of kBsThreads keys in input order: a
// key's place = its digit's first place + the places earlier rounds took + the
// keys of its digit in lower waves of this round + its rank in its wave (eight
// ballots of the digit's bits). The pointers are LDS or global.
template <class SK, class SV, class DK, class DV>
__device__ void rank_pass(SK sk, SV sv, DK dk, DV dv, int m, int shift, RankLds &R)
{
    const int t = threadIdx.x, w = t >> 6;
    if (t < kRsBins) R.tot[t] = 0;
    for (int x = t; x < kBsWaves * kRsBins; x += kBsThreads) (&R.cnt[0][0])[x] = 0;
    __syncthreads();
    for (int j = t; j < m; j += kBsThreads) atomicAdd(&R.tot[(unsigned)(sk[j] >> shift) & 255u], 1u);
    __syncthreads();
    const unsigned ls = scan256_excl(t < kRsBins ? R.tot[t] : 0u, R.wsum);
    if (t < kRsBins) {
        R.lstart[t] = ls;
        R.run[t] = 0;
    }
    __syncthreads();
    const int lane = t & 63;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int r0 = 0; r0 < m; r0 += kBsThreads) {
        const int j = r0 + t;
        const bool valid = j < m;
        const unsigned long long k = valid ? sk[j] : 0ull;
        const int v = valid ? sv[j] : 0;
        const unsigned dg = (unsigned)(k >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bal = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bal : ~bal;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) R.cnt[w][dg] = (unsigned)__popcll(peers);
        __syncthreads();
        if (t < kRsBins) {
            unsigned acc = 0;
#pragma unroll
            for (int ww = 0; ww < kBsWaves; ++ww) {
                const unsigned c = R.cnt[ww][t];
                R.cnt[ww][t] = acc;
                acc += c;
            }
            R.tot[t] = acc;
        }
        __syncthreads();
        if (valid) {
            const unsigned at = R.lstart[dg] + R.run[dg] + R.cnt[w][dg] + (unsigned)rank;
            dk[at] = k;
            dv[at] = v;
        }
        __syncthreads();
        if (t < kRsBins) {
            R.run[t] += R.tot[t];
#pragma unroll
            for (int ww = 0; ww < kBsWaves; ++ww) R.cnt[ww][t] = 0;
        }
        __syncthreads();
    }
}

// the highest key bit that differs inside key[0 .. m-1] (-1: all equal)
template <class SK>
__device__ int varying_top_bit(SK key, int m, unsigned long long *red)
{
    const int t = threadIdx.x;
    if (t == 0) *red = 0ull;
    __syncthreads();
    const unsigned long long k0 = key[0];
    unsigned long long x = 0;
    for (int j = t; j < m; j += kBsThreads) x |= key[j] ^ k0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    if ((t & 63) == 0 && x) atomicOr(red, x);
    __syncthreads();
    const unsigned long long all = *red;
    __syncthreads();
    return all ? 63 - __clzll((long long)all) : -1;
}

// Bucket blockIdx.x of the S + 1 (blocks of one wave), in place in (key,
// val): lane l holds elements r * 64 + l (r < kSortPer) in registers. A pass
// by the 8 bits at `shift`: digit counts (LDS atomics) and their exclusive
// scan give each digit's first place; then row by row, in element order, a
// key's place is its digit's next free place plus its rank among the lanes
// of the row with its digit (eight ballots); the row's lanes move their
// digit's next place on. The keys go to those places in LDS and come back to
// the registers in element order. A bucket above kBucketCap is listed in
// big[1 ..] (big[0] counts them) for k_bucket_sort_big.
constexpr int kSortPer = kBucketCap / 64;

__global__ __launch_bounds__(64) void k_bucket_sort(const Ctl *ctl, unsigned long long *__restrict__ key,
                                                    int *__restrict__ val, const unsigned *__restrict__ bstart,
                                                    int *__restrict__ big)
{
    __shared__ unsigned long long sk[kBucketCap];
    __shared__ int sv[kBucketCap];
    __shared__ unsigned next[kRsBins];
    if (ctl->error) return;
    const int b = blockIdx.x, lane = threadIdx.x;
    const int base = (int)bstart[b], m = (int)bstart[b + 1] - base;
    if (m <= 1) return;
    if (m > kBucketCap) {
        if (lane == 0) big[1 + atomicAdd(&big[0], 1)] = b;
        return;
    }
    unsigned long long k[kSortPer];
    int v[kSortPer];
    unsigned long long diff = 0;
    const unsigned long long k0 = key[base];
#pragma unroll
    for (int r = 0; r < kSortPer; ++r) {
        const int idx = r * 64 + lane;
        k[r] = idx < m ? key[base + idx] : k0;
        v[r] = idx < m ? val[base + idx] : 0;
        diff |= k[r] ^ k0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) diff |= __shfl_xor(diff, o);
    const int top = diff ? 63 - __clzll((long long)diff) : -1;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int shift = 0; shift <= top; shift += 8) {
#pragma unroll
        for (int q = 0; q < kRsBins / 64; ++q) next[q * 64 + lane] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortPer; ++r) {
            if (r * 64 >= m) break;
            if (r * 64 + lane < m) atomicAdd(&next[(unsigned)(k[r] >> shift) & 255u], 1u);
        }
        __syncthreads();
        {
            // lane l scans digits 4l .. 4l+3
            unsigned c[4], sum = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c[q] = next[lane * 4 + q];
                sum += c[q];
            }
            unsigned x = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned y = __shfl_up(x, o);
                if (lane >= o) x += y;
            }
            unsigned run = x - sum;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                next[lane * 4 + q] = run;
                run += c[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortPer; ++r) {
            if (r * 64 >= m) break; // rows past the bucket: no key
            const bool valid = r * 64 + lane < m;
            const unsigned dg = (unsigned)(k[r] >> shift) & 255u;
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit) {
                const unsigned long long bal = __ballot((dg >> bit) & 1u);
                peers &= ((dg >> bit) & 1u) ? bal : ~bal;
            }
            const unsigned rank = (unsigned)__popcll(peers & below);
            const unsigned at = valid ? next[dg] + rank : 0u;
            __syncthreads();
            if (valid) {
                if (rank == 0) next[dg] += (unsigned)__popcll(peers);
                sk[at] = k[r];
                sv[at] = v[r];
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < kSortPer; ++r) {
            const int idx = r * 64 + lane;
            if (idx < m) {
                k[r] = sk[idx];
                v[r] = sv[idx];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kSortPer; ++r) {
        const int idx = r * 64 + lane;
        if (idx < m) {
            key[base + idx] = k[r];
            val[base + idx] = v[r];
        }
    }
}

// The listed buckets (more than kBucketCap bodies), one block each, in place
// in (key, val) through (tkey, tval), free n-element buffers.
constexpr int kBigBlocks = 64;

__global__ __launch_bounds__(kBsThreads) void k_bucket_sort_big(Ctl *ctl, unsigned long long *__restrict__ key,
                                                                int *__restrict__ val, unsigned long long *__restrict__ tkey,
                                                                int *__restrict__ tval, const unsigned *__restrict__ bstart,
                                                                const int *__restrict__ big)
{
    __shared__ RankLds R;
    __shared__ unsigned long long red;
    if (ctl->error) return;
    const int t = threadIdx.x, nbig = big[0];
    for (int q = blockIdx.x; q < nbig; q += gridDim.x) {
        const int b = big[1 + q];
        const int base = (int)bstart[b], m = (int)bstart[b + 1] - base;
        if (t == 0) atomicAdd(&ctl->large_buckets, 1);
        unsigned long long *k0 = key + base, *k1 = tkey + base;
        int *v0 = val + base, *v1 = tval + base;
        const int top = varying_top_bit(k0, m, &red);
        bool in_tmp = false;
        for (int shift = 0; shift <= top; shift += 8) {
            if (in_tmp) rank_pass(k1, v1, k0, v0, m, shift, R);
            else rank_pass(k0, v0, k1, v1, m, shift, R);
            in_tmp = !in_tmp;
        }
        if (in_tmp) {
            for (int j = t; j < m; j += kBsThreads) {
                k0[j] = k1[j];
                v0[j] = v1[j];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Runs of equal 21-level keys: their bodies' paths extended to 84 levels
// (deep[0..2][pos], levels 21 .. 83), each run ordered by them, one thread
// per run.
// ---------------------------------------------------------------------------
// levels 21 .. 83 of body p's path, three words of 21 levels
__device__ __forceinline__ void deep_words(float rx, float ry, float rz, float rr, float4 p,
                                           unsigned long long *out)
{
    float cx = rx, cy = ry, cz = rz, r = rr;
    for (int L = 0; L < kKeyLevels; ++L) key_descend(cx, cy, cz, r, key_octant(cx, cy, cz, p));
    for (int w = 0; w < kDeepWords; ++w) {
        unsigned long long x = 0;
        for (int L = 0; L < kKeyLevels; ++L) {
            const int o = key_octant(cx, cy, cz, p);
            x = (x << 3) | (unsigned long long)o;
            key_descend(cx, cy, cz, r, o);
        }
        out[w] = x;
    }
}

// deep path of run element x before that of y (words at base[w * n + .])
__device__ __forceinline__ bool deep_less(const unsigned long long *base, size_t n, int x, int y)
{
    for (int w = 0; w < kDeepWords; ++w) {
        const unsigned long long a = base[w * n + x], b = base[w * n + y];
        if (a != b) return a < b;
    }
    return false;
}

// A run longer than kMaxRun (bodies i .. i+m-1): ordered through global
// memory by one thread. The deep words go to `deep` in the run's current
// order, a permutation of the run is merge-sorted (stable, in perm with tmp
// as the merge buffer), then the bodies and their deep words are gathered
// through staging copies. perm, tmp, stage_k and stage_v are n-element
// buffers that nothing else uses at this point; each run works in its own
// range [i, i+m) of them.
__device__ void order_long_run(const Dev &d, int i, int m, int *val, unsigned long long *deep, int *perm, int *tmp,
                               unsigned long long *stage_k, int *stage_v)
{
    const size_t n = (size_t)d.n;
    const Ctl *ctl = d.ctl;
    const float rx = ctl->root[0], ry = ctl->root[1], rz = ctl->root[2], rr = ctl->radius;
    for (int a = 0; a < m; ++a) {
        unsigned long long w3[kDeepWords];
        deep_words(rx, ry, rz, rr, d.pm4[val[i + a]], w3);
        for (int w = 0; w < kDeepWords; ++w) deep[w * n + i + a] = w3[w];
        perm[i + a] = a;
    }
    const unsigned long long *base = deep + i;
    int *src = perm + i, *dst = tmp + i;
    for (int width = 1; width < m; width <<= 1) {
        for (int lo = 0; lo < m; lo += 2 * width) {
            const int mid = min(lo + width, m), hi = min(lo + 2 * width, m);
            int a = lo, b = mid, o = lo;
            while (a < mid && b < hi) dst[o++] = deep_less(base, n, src[b], src[a]) ? src[b++] : src[a++];
            while (a < mid) dst[o++] = src[a++];
            while (b < hi) dst[o++] = src[b++];
        }
        int *t = src;
        src = dst;
        dst = t;
    }
    for (int a = 0; a < m; ++a) stage_v[i + a] = val[i + src[a]];
    for (int a = 0; a < m; ++a) val[i + a] = stage_v[i + a];
    for (int w = 0; w < kDeepWords; ++w) {
        for (int a = 0; a < m; ++a) stage_k[i + a] = deep[w * n + i + src[a]];
        for (int a = 0; a < m; ++a) deep[w * n + i + a] = stage_k[i + a];
    }
    for (int a = 1; a < m; ++a)
        if (!deep_less(base, n, a - 1, a)) flag_error(d.ctl, kErrDepth); // one 84-level path
}

__global__ __launch_bounds__(256) void k_deep_runs(Dev d, const unsigned long long *__restrict__ key, int *val,
                                                   unsigned long long *__restrict__ deep, int *perm, int *tmp,
                                                   unsigned long long *stage_k, int *stage_v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (i + 1 >= n || ctl->error) return;
    const unsigned long long k = key[i];
    if (key[i + 1] != k || (i > 0 && key[i - 1] == k)) return; // not the start of a run
    int m = 2;
    while (i + m < n && key[i + m] == k) ++m;
    if (m > kMaxRun) {
        order_long_run(d, i, m, val, deep, perm, tmp, stage_k, stage_v);
        return;
    }
    unsigned long long dk[kMaxRun][kDeepWords];
    int vv[kMaxRun];
    const float rx = ctl->root[0], ry = ctl->root[1], rz = ctl->root[2], rr = ctl->radius;
    for (int a = 0; a < m; ++a) {
        vv[a] = val[i + a];
        deep_words(rx, ry, rz, rr, d.pm4[vv[a]], dk[a]);
    }
    // stable insertion sort by the deep words
    for (int a = 1; a < m; ++a) {
        unsigned long long c0 = dk[a][0], c1 = dk[a][1], c2 = dk[a][2];
        const int cv = vv[a];
        int b = a - 1;
        while (b >= 0 && (dk[b][0] > c0 || (dk[b][0] == c0 && (dk[b][1] > c1 || (dk[b][1] == c1 && dk[b][2] > c2))))) {
            dk[b + 1][0] = dk[b][0];
            dk[b + 1][1] = dk[b][1];
            dk[b + 1][2] = dk[b][2];
            vv[b + 1] = vv[b];
            --b;
        }
        dk[b + 1][0] = c0;
        dk[b + 1][1] = c1;
        dk[b + 1][2] = c2;
        vv[b + 1] = cv;
    }
    for (int a = 0; a < m; ++a) {
        // two bodies on one 84-level path are closer than the tree can separate
        if (a > 0 && dk[a][0] == dk[a - 1][0] && dk[a][1] == dk[a - 1][1] && dk[a][2] == dk[a - 1][2])
            flag_error(ctl, kErrDepth);
        val[i + a] = vv[a];
        for (int w = 0; w < kDeepWords; ++w) deep[(size_t)w * n + i + a] = dk[a][w];
    }
}

// Common prefix in bits (3 per level) of sorted bodies i and j, whose keys are
// ki and kj.
__device__ __forceinline__ int delta_of(unsigned long long ki, unsigned long long kj,
                                        const unsigned long long *__restrict__ deep, int n, int i, int j)
{
    const unsigned long long x = ki ^ kj;
    if (x) return __clzll((long long)x) - 1;
    // equal 21-level keys: i and j are in one run, whose deep words are set
#pragma unroll
    for (int w = 0; w < kDeepWords; ++w) {
        const unsigned long long y = deep[(size_t)w * n + i] ^ deep[(size_t)w * n + j];
        if (y) return 63 * (w + 1) + __clzll((long long)y) - 1;
    }
    return 63 * (kDeepWords + 1);
}

// Common prefix of sorted keys i and j in bits (3 per level); -1 outside.
__device__ __forceinline__ int key_delta(const unsigned long long *__restrict__ key,
                                         const unsigned long long *__restrict__ deep, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    return delta_of(key[i], key[j], deep, n, i, j);
}

// octant of sorted body `pos` at `level`
__device__ __forceinline__ int key_digit(const unsigned long long *__restrict__ key,
                                         const unsigned long long *__restrict__ deep, int n, int pos, int level)
{
    const int w = level / kKeyLevels, L = level % kKeyLevels;
    const unsigned long long x = w == 0 ? key[pos] : deep[(size_t)(w - 1) * n + pos];
    return (int)((x >> (3 * (kKeyLevels - 1 - L))) & 7ull);
}

// ---------------------------------------------------------------------------
// Binary radix tree over the sorted keys (Karras 2012, fig. 4), internal node
// i of n - 1 (node 0 is the root): its range [lo, hi] and split, the parent
// links of its two children (internal or body), and the octree levels it
// holds: ktop .. ktop + own - 1, the levels k with 3k in (delta(parent),
// delta(i)]. delta(parent) is the larger of the prefixes the range shares with
// its two outside neighbours (the parent extends the range towards the
// neighbour that shares more; -1 for the root).
//
// The searches of a block's nodes read the keys from a window in LDS (the
// block's own keys and kBtHalo on each side, loaded once, coalesced); only a
// probe past the window (a node whose range reaches further) reads the key
// from device memory. Most ranges are short, so the searches' dependent loads
// become LDS reads.
// ---------------------------------------------------------------------------
constexpr int kBtThreads = 512;
constexpr int kBtHalo = 512;
constexpr int kBtSpan = kBtThreads + 2 * kBtHalo;

// keys through the block's LDS window where it has them, else from device memory
struct KeysAny {
    const unsigned long long *ks;
    const unsigned long long *__restrict__ key;
    int base;
    __device__ __forceinline__ unsigned long long operator()(int j) const
    {
        const unsigned t = (unsigned)(j - base);
        return t < (unsigned)kBtSpan ? ks[t] : key[j];
    }
};

// delta(i, j) > dd for i's key ki, with lim = limit(dd). Keys use 63 bits, so
// delta = clz(ki ^ kj) - 1 below 63, and delta > dd (dd < 63) holds exactly
// when ki ^ kj < 2^(62 - dd) (equal keys included: their delta is >= 63); one
// 64-bit compare instead of the count of leading zeros. A dd of 63 or more (i
// inside a run of equal keys) compares the deep words.
__device__ __forceinline__ unsigned long long delta_limit(int dd) { return dd < 63 ? 1ull << (62 - dd) : 0ull; }

template <class K>
__device__ __forceinline__ bool delta_above(const K &ka, const unsigned long long *__restrict__ deep, int n, int i,
                                            unsigned long long ki, int j, int dd, unsigned long long lim)
{
    if (j < 0 || j >= n) return false;
    if (lim) return (ki ^ ka(j)) < lim;
    return delta_of(ki, ka(j), deep, n, i, j) > dd;
}

template <class K>
__device__ __forceinline__ int delta_at(const K &ka, const unsigned long long *__restrict__ deep, int n, int i,
                                        unsigned long long ki, int j)
{
    return j < 0 || j >= n ? -1 : delta_of(ki, ka(j), deep, n, i, j);
}

struct BinOut {
    int *__restrict__ iparent, *__restrict__ lparent, *__restrict__ lo, *__restrict__ own, *__restrict__ ktop;
};

// Internal node i: its searches, then its outputs. dir and dmin come from
// the neighbours i - 1 and i + 1. n <= 2^26 (the pool limit), so every probe
// fits an int.
template <class K>
__device__ __forceinline__ void bintree_node(const Dev &d, const K &ka, const unsigned long long *__restrict__ deep,
                                             int n, int i, const BinOut &o)
{
    const unsigned long long ki = ka(i);
    const int dir = delta_at(ka, deep, n, i, ki, i + 1) - delta_at(ka, deep, n, i, ki, i - 1) > 0 ? 1 : -1;
    const int dmin = delta_at(ka, deep, n, i, ki, i - dir);
    const unsigned long long lmin = delta_limit(dmin);
    int lmax = 2;
    while (delta_above(ka, deep, n, i, ki, i + lmax * dir, dmin, lmin)) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta_above(ka, deep, n, i, ki, i + (l + t) * dir, dmin, lmin)) l += t;
    const int j = i + l * dir;
    const int dn = delta_at(ka, deep, n, i, ki, j);
    const unsigned long long ln = delta_limit(dn);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta_above(ka, deep, n, i, ki, i + (s + t) * dir, dn, ln)) s += t;
    } while (t > 1);
    const int g = i + s * dir + min(dir, 0);
    const int lo = min(i, j), hi = max(i, j);
    if (lo == g) o.lparent[g] = i;
    else o.iparent[g] = i;
    if (hi == g + 1) o.lparent[g + 1] = i;
    else o.iparent[g + 1] = i;
    o.lo[i] = lo;
    if (i == 0) o.iparent[0] = -1;
    const int dp = max(delta_at(ka, deep, n, lo, ka(lo), lo - 1), delta_at(ka, deep, n, hi, ka(hi), hi + 1));
    const int kt = dp < 0 ? 0 : dp / 3 + 1;
    int kb = dn / 3;
    if (kb >= kBuildLevelLimit) {
        flag_error(d.ctl, kErrDepth);
        kb = kBuildLevelLimit - 1;
    }
    o.own[i] = max(0, kb - kt + 1);
    o.ktop[i] = kt;
}

__global__ __launch_bounds__(kBtThreads) void k_bintree(Dev d, const unsigned long long *__restrict__ key,
                                                        const unsigned long long *__restrict__ deep, BinOut o)
{
    __shared__ unsigned long long ks[kBtSpan];
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (ctl->error) return; // block-uniform
    const int base = (int)blockIdx.x * kBtThreads - kBtHalo;
    for (int t = threadIdx.x; t < kBtSpan; t += kBtThreads) {
        const int g = base + t;
        ks[t] = g >= 0 && g < n ? key[g] : 0ull;
    }
    __syncthreads();
    const int i = (int)blockIdx.x * kBtThreads + (int)threadIdx.x;
    if (i >= n - 1) return;
    bintree_node(d, KeysAny{ks, key, base}, deep, n, i, o);
}

// ---------------------------------------------------------------------------
// Exclusive scan of the cells each internal node holds (-> its first cell):
// tiles of kScanTile scanned with their sums kept (k_scan_tiles), the sums
// scanned by one block (k_scan_sums); readers add the two (first_cell).
// The same passes find the tree's depth: the deepest cell's level + 1 (its
// children are bodies), per tile and then over the tiles, into the running
// maximum (buildtree.cl:199) with one atomic.
// ---------------------------------------------------------------------------
constexpr int kScanPer = 4; // one int4 per thread: coalesced
constexpr int kScanTile = 256 * kScanPer;
static_assert(kScanTile == kKeyScanTile, "the engine sizes the tile sums with kKeyScanTile");

__device__ __forceinline__ int4 load4(const int *__restrict__ a, int i0, int m)
{
    if (i0 + 3 < m) return *reinterpret_cast<const int4 *>(a + i0);
    return make_int4(i0 < m ? a[i0] : 0, i0 + 1 < m ? a[i0 + 1] : 0, i0 + 2 < m ? a[i0 + 2] : 0, 0);
}

__global__ __launch_bounds__(256) void k_scan_tiles(const Ctl *ctl, const int *__restrict__ in,
                                                    const int *__restrict__ ktop, int *__restrict__ out, int m,
                                                    int *__restrict__ sums, int *__restrict__ tdepth)
{
    __shared__ int wsum[4], wdep[4];
    if (ctl->error) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int i0 = blockIdx.x * kScanTile + t * kScanPer;
    const int4 v = load4(in, i0, m), kt = load4(ktop, i0, m);
    const int acc = v.x + v.y + v.z + v.w;
    int dep = 0;
    if (v.x > 0) dep = max(dep, kt.x + v.x);
    if (v.y > 0) dep = max(dep, kt.y + v.y);
    if (v.z > 0) dep = max(dep, kt.z + v.z);
    if (v.w > 0) dep = max(dep, kt.w + v.w);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dep = max(dep, __shfl_xor(dep, o));
    if (lane == 0) wdep[w] = dep;
    int x = acc; // inclusive scan of the per-thread totals over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int run = x - acc;
    for (int ww = 0; ww < w; ++ww) run += wsum[ww];
    const int4 e = make_int4(run, run + v.x, run + v.x + v.y, run + v.x + v.y + v.z);
    if (i0 + 3 < m) {
        *reinterpret_cast<int4 *>(out + i0) = e;
    } else {
        if (i0 < m) out[i0] = e.x;
        if (i0 + 1 < m) out[i0 + 1] = e.y;
        if (i0 + 2 < m) out[i0 + 2] = e.z;
    }
    if (t == 0) {
        sums[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        tdepth[blockIdx.x] = max(max(wdep[0], wdep[1]), max(wdep[2], wdep[3]));
    }
}

// one block: exclusive scan of the tile sums in place (each thread a run of
// them); the tiles' depths into the running maximum
__global__ __launch_bounds__(1024) void k_scan_sums(Dev d, int *__restrict__ sums, int ntiles,
                                                   const int *__restrict__ tdepth)
{
    __shared__ int s[1024];
    __shared__ int wdep[16];
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const int t = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    int acc = 0, dep = 0;
    for (int k = 0; k < per; ++k) {
        const int j = t * per + k;
        if (j < ntiles) {
            acc += sums[j];
            dep = max(dep, tdepth[j]);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dep = max(dep, __shfl_xor(dep, o));
    if ((t & 63) == 0) wdep[t >> 6] = dep;
    s[t] = acc;
    __syncthreads();
    for (int st = 1; st < 1024; st <<= 1) {
        const int u = t >= st ? s[t - st] : 0;
        __syncthreads();
        s[t] += u;
        __syncthreads();
    }
    int run = s[t] - acc;
    for (int k = 0; k < per; ++k) {
        const int j = t * per + k;
        if (j < ntiles) {
            const int c = sums[j];
            sums[j] = run;
            run += c;
        }
    }
    if (t == 0) {
        for (int k = 0; k < 16; ++k) dep = max(dep, wdep[k]);
        if (dep > 0) __hip_atomic_fetch_max(&ctl->max_depth, dep, __ATOMIC_RELAXED, AGENT);
    }
    // the cell count (the total of the scan), the pool check, and the root's record
    if (t == 1023) {
        const int C = s[1023];
        if (C > d.cap) {
            flag_error(ctl, kErrPool);
        } else {
            ctl->n_cells = C;
            d.cinfo[0] = make_int4(-1, 0, (int)ctl->epoch, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Cells and slots. Cell first_cell(i) + (k - ktop[i]) is internal node i's cell
// of level k; the root (level 0, held by binary node 0) is cell 0.
// ---------------------------------------------------------------------------
struct KeyTree {
    const unsigned long long *key, *deep;
    const int *val, *iparent, *lparent, *lo, *own, *ktop, *obase, *tsum;
};

__device__ __forceinline__ int first_cell(const KeyTree &kt, int i) { return kt.obase[i] + kt.tsum[i / kScanTile]; }

// The deepest cell of the nearest binary ancestor (p itself included) that
// holds a cell, or -1 (an inconsistent tree; the caller flags it). A node that
// holds none shares its level with its parent, so the walk is at most a few
// steps; it is bounded anyway.
__device__ __forceinline__ int holding_cell(const KeyTree &kt, int p)
{
    for (int steps = 0; steps < 4 * kBuildLevelLimit * 3 && p >= 0; ++steps) {
        const int o = kt.own[p];
        if (o != 0) return first_cell(kt, p) + o - 1;
        p = kt.iparent[p];
    }
    return -1;
}

// one body: the root is the one cell (k_scan_sums counts the cells otherwise)
__global__ void k_one_cell(Dev d)
{
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    ctl->n_cells = 1;
    d.cinfo[0] = make_int4(-1, 0, (int)ctl->epoch, 0);
}

// Position g: the cells of internal node g (g < n - 1) -- each in the slot of
// its octant in its parent, and its record -- and sorted body g, in the
// deepest cell of its nearest holding ancestor, at the level one below the
// larger prefix it shares with a neighbour. The kernel is bound by the
// latency of its dependent reads, not by its arithmetic: a thread takes
// kLinkPer positions (kLinkPer * 256 apart in a block's span), and their
// loads are issued together, in two rounds of independent reads (clamped
// indices, no branch in between).
constexpr int kLinkPer = 2;
constexpr int kLinkSpan = 256 * kLinkPer;

__device__ __forceinline__ int deepest_of(const KeyTree &kt, int own, int ob, int ts, int p)
{
    // the holding cell of binary node p, given p's own, obase and tile sum:
    // p itself when it holds cells, else the walk up from its parent
    return own != 0 ? ob + ts + own - 1 : (p >= 0 ? holding_cell(kt, kt.iparent[p]) : -1);
}

__global__ __launch_bounds__(256) void k_link(Dev d, KeyTree kt)
{
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const uint32_t ep = ctl->epoch;
    const int g0 = blockIdx.x * kLinkSpan + threadIdx.x;
    if (n < 2) { // the root holds the one body
        if (g0 == 0) d.slot[key_digit(kt.key, kt.deep, n, 0, 0)] = slot_tag(ep, kt.val[0]);
        return;
    }
    // round 1: each position's node and leaf records and the keys around it
    int own[kLinkPer], top[kLinkPer], lo[kLinkPer], ip[kLinkPer], ob[kLinkPer], ts[kLinkPer], lp[kLinkPer];
    int body[kLinkPer];
    unsigned long long kg[kLinkPer], km[kLinkPer], kq[kLinkPer];
#pragma unroll
    for (int u = 0; u < kLinkPer; ++u) {
        const int g = min(g0 + u * 256, n - 1);
        const int gi = min(g, n - 2);
        own[u] = kt.own[gi];
        top[u] = kt.ktop[gi];
        lo[u] = kt.lo[gi];
        ip[u] = kt.iparent[gi];
        ob[u] = kt.obase[gi];
        ts[u] = kt.tsum[gi / kScanTile];
        lp[u] = kt.lparent[g];
        body[u] = kt.val[g];
        kg[u] = kt.key[g];
        km[u] = kt.key[max(g - 1, 0)];
        kq[u] = kt.key[min(g + 1, n - 1)];
    }
    // round 2: the parents' records, and the key of each node's first body
    int own_ip[kLinkPer], ob_ip[kLinkPer], ts_ip[kLinkPer], own_lp[kLinkPer], ob_lp[kLinkPer], ts_lp[kLinkPer];
    unsigned long long klo[kLinkPer];
#pragma unroll
    for (int u = 0; u < kLinkPer; ++u) {
        const int ipc = max(ip[u], 0);
        own_ip[u] = kt.own[ipc];
        ob_ip[u] = kt.obase[ipc];
        ts_ip[u] = kt.tsum[ipc / kScanTile];
        own_lp[u] = kt.own[lp[u]];
        ob_lp[u] = kt.obase[lp[u]];
        ts_lp[u] = kt.tsum[lp[u] / kScanTile];
        klo[u] = kt.key[lo[u]];
    }
#pragma unroll
    for (int u = 0; u < kLinkPer; ++u) {
        const int g = g0 + u * 256;
        if (g >= n) break;
        if (g < n - 1 && own[u] > 0) {
            int par = -1;
            if (top[u] > 0) {
                par = deepest_of(kt, own_ip[u], ob_ip[u], ts_ip[u], ip[u]);
                if (par < 0) {
                    flag_error(ctl, kErrSync);
                    return;
                }
            }
            const int base = ob[u] + ts[u];
            for (int k = top[u]; k < top[u] + own[u]; ++k) {
                const int cid = base + (k - top[u]);
                if (k > 0) {
                    const int L = k - 1;
                    const int dg = L < kKeyLevels ? (int)((klo[u] >> (3 * (kKeyLevels - 1 - L))) & 7ull)
                                                  : key_digit(kt.key, kt.deep, n, lo[u], L);
                    d.slot[(size_t)par * 8 + dg] = slot_tag(ep, n + cid);
                    d.cinfo[cid] = make_int4(par, k | (dg << 8), (int)ep, 0);
                }
                par = cid;
            }
        }
        const int par = deepest_of(kt, own_lp[u], ob_lp[u], ts_lp[u], lp[u]);
        if (par < 0) {
            flag_error(ctl, kErrSync);
            return;
        }
        // common prefixes with the neighbours from the keys in hand (the deep
        // words only when a key is repeated)
        const int dm = g == 0 ? -1
                              : (kg[u] != km[u] ? __clzll((long long)(kg[u] ^ km[u])) - 1
                                                : key_delta(kt.key, kt.deep, n, g, g - 1));
        const int dq = g == n - 1 ? -1
                                  : (kg[u] != kq[u] ? __clzll((long long)(kg[u] ^ kq[u])) - 1
                                                    : key_delta(kt.key, kt.deep, n, g, g + 1));
        const int lv = max(dm, dq) / 3;
        const int dg = lv < kKeyLevels ? (int)((kg[u] >> (3 * (kKeyLevels - 1 - lv))) & 7ull)
                                       : key_digit(kt.key, kt.deep, n, g, lv);
        d.slot[(size_t)par * 8 + dg] = slot_tag(ep, body[u]);
    }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
hipError_t launch_build_keys(const Dev &d, bool warm, bool staged, hipStream_t st)
{
    const int n = d.n;
    const dim3 b256(256), g256((n + 255) / 256);
    unsigned long long *ka = d.kb_key[0], *kb = d.kb_key[1];
    int *va = d.kb_val[0], *vb = d.kb_val[1];
    if (warm) {
        // keys in the last tree's order into [1], splitters from the last
        // tree's sorted keys in [0], bucketed into [0], each bucket sorted in
        // place ([1] is the spill space of large buckets)
        // (the buckets of the bodies go to kb_lo, which the binary tree fills later)
        const int S = bucket_samples(n);
        const dim3 gw((n + kWinBodies - 1) / kWinBodies);
        hipLaunchKernelGGL(k_splitters, dim3((S + 1 + 255) / 256), b256, 0, st, d.ctl, ka, S, d.kb_spl, d.kb_bcnt,
                           d.kb_bfill, d.kb_big);
        hipLaunchKernelGGL(k_keys_warm, dim3((n + kWarmBodies - 1) / kWarmBodies), b256, 0, st, d, d.sorted,
                           staged ? d.bnode : nullptr, d.kb_spl, S, kb, vb, d.kb_lo, d.kb_bcnt);
        hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, st, d.ctl, d.kb_bcnt, S, d.kb_bstart);
        hipLaunchKernelGGL(k_bucket_scatter, gw, b256, 0, st, d.ctl, kb, vb, d.kb_lo, ka, va, n, d.kb_bstart,
                           d.kb_bfill);
        hipLaunchKernelGGL(k_bucket_sort, dim3(S + 1), dim3(64), 0, st, d.ctl, ka, va, d.kb_bstart, d.kb_big);
        hipLaunchKernelGGL(k_bucket_sort_big, dim3(kBigBlocks), dim3(kBsThreads), 0, st, d.ctl, ka, va, kb, vb,
                           d.kb_bstart, d.kb_big);
    } else {
        const int ntiles = (n + kRsTile - 1) / kRsTile;
        hipLaunchKernelGGL(k_keys, g256, b256, 0, st, d, nullptr, ka, va);
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 8 * pass;
            hipLaunchKernelGGL(k_rs_count, dim3(ntiles), dim3(kRsThreads), 0, st, ka, n, shift, d.kb_hist, ntiles);
            hipLaunchKernelGGL(k_rs_offsets, dim3(kRsBins), dim3(1024), 0, st, d.kb_hist, ntiles, d.kb_offs, d.kb_tot);
            hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, st, ka, va, kb, vb, n, shift, d.kb_hist,
                               d.kb_offs, d.kb_tot, ntiles);
            unsigned long long *tk = ka;
            ka = kb;
            kb = tk;
            int *tv = va;
            va = vb;
            vb = tv;
        }
        // eight passes: the sorted keys are back in kb_key[0] / kb_val[0]
    }
    // the other pair and the binary tree's arrays (written later) serve long runs
    hipLaunchKernelGGL(k_deep_runs, g256, b256, 0, st, d, ka, va, d.kb_deep, d.kb_ipar, d.kb_lpar, kb, vb);
    KeyTree kt{ka, d.kb_deep, va, d.kb_ipar, d.kb_lpar, d.kb_lo, d.kb_own, d.kb_ktop, d.kb_obase, d.kb_scan};
    if (n >= 2) {
        const dim3 gi((n - 1 + kBtThreads - 1) / kBtThreads);
        const BinOut bo{d.kb_ipar, d.kb_lpar, d.kb_lo, d.kb_own, d.kb_ktop};
        hipLaunchKernelGGL(k_bintree, gi, dim3(kBtThreads), 0, st, d, ka, d.kb_deep, bo);
        const int m = n - 1, nt = (m + kScanTile - 1) / kScanTile;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, st, d.ctl, d.kb_own, d.kb_ktop, d.kb_obase, m,
                           d.kb_scan, d.kb_tdep);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, d, d.kb_scan, nt, d.kb_tdep);
    } else {
        hipLaunchKernelGGL(k_one_cell, dim3(1), dim3(1), 0, st, d);
    }
    hipLaunchKernelGGL(k_link, dim3((n + kLinkSpan - 1) / kLinkSpan), b256, 0, st, d, kt);
    return hipGetLastError();
}

} // namespace nb
