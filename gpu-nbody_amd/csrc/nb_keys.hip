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
// Keys: 21 levels of each body's octant path, level 0 in bits 62..60.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_keys(Dev d, unsigned long long *key, int *val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const Ctl *ctl = d.ctl;
    if (i >= d.n || ctl->error) return;
    const float4 p = d.pm4[i];
    float cx = ctl->root[0], cy = ctl->root[1], cz = ctl->root[2], r = ctl->radius;
    unsigned long long k = 0;
#pragma unroll
    for (int L = 0; L < kKeyLevels; ++L) {
        const int o = key_octant(cx, cy, cz, p);
        k = (k << 3) | (unsigned long long)o;
        key_descend(cx, cy, cz, r, o);
    }
    key[i] = k;
    val[i] = i;
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

// Common prefix of sorted keys i and j in bits (3 per level); -1 outside.
__device__ __forceinline__ int key_delta(const unsigned long long *__restrict__ key,
                                         const unsigned long long *__restrict__ deep, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const unsigned long long x = key[i] ^ key[j];
    if (x) return __clzll((long long)x) - 1;
    // equal 21-level keys: i and j are in one run, whose deep words are set
#pragma unroll
    for (int w = 0; w < kDeepWords; ++w) {
        const unsigned long long y = deep[(size_t)w * n + i] ^ deep[(size_t)w * n + j];
        if (y) return 63 * (w + 1) + __clzll((long long)y) - 1;
    }
    return 63 * (kDeepWords + 1);
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
// i of n - 1 (node 0 is the root): its range, the common prefix of its keys
// in bits, and the parent links of its two children (internal or body).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bintree(Dev d, const unsigned long long *__restrict__ key,
                                                 const unsigned long long *__restrict__ deep, int *__restrict__ iparent,
                                                 int *__restrict__ lparent, int *__restrict__ dnode,
                                                 int *__restrict__ lo_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = d.n;
    if (i >= n - 1 || d.ctl->error) return;
    const int dir = key_delta(key, deep, n, i, i + 1) - key_delta(key, deep, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = key_delta(key, deep, n, i, i - dir);
    long long lmax = 2;
    while (key_delta(key, deep, n, i, (int)(i + lmax * dir)) > dmin) lmax <<= 1;
    long long l = 0;
    for (long long t = lmax >> 1; t >= 1; t >>= 1)
        if (key_delta(key, deep, n, i, (int)(i + (l + t) * dir)) > dmin) l += t;
    const int j = (int)(i + l * dir);
    const int dn = key_delta(key, deep, n, i, j);
    long long s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (key_delta(key, deep, n, i, (int)(i + (s + t) * dir)) > dn) s += t;
    } while (t > 1);
    const int g = (int)(i + s * dir + min(dir, 0));
    const int lo = min(i, j), hi = max(i, j);
    if (lo == g) lparent[g] = i;
    else iparent[g] = i;
    if (hi == g + 1) lparent[g + 1] = i;
    else iparent[g + 1] = i;
    dnode[i] = dn;
    lo_out[i] = lo;
    if (i == 0) iparent[0] = -1;
}

// Octree levels held by internal node i: ktop .. ktop + own - 1, the levels k
// with 3k in (delta(parent), delta(i)] (from level 0 at the root).
__global__ __launch_bounds__(256) void k_own(Dev d, const int *__restrict__ iparent, const int *__restrict__ dnode,
                                             int *__restrict__ own, int *__restrict__ ktop)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = d.n;
    if (i >= n - 1 || d.ctl->error) return; // no binary tree was built
    const int kt = i == 0 ? 0 : dnode[iparent[i]] / 3 + 1;
    int kb = dnode[i] / 3;
    if (kb >= kBuildLevelLimit) {
        flag_error(d.ctl, kErrDepth);
        kb = kBuildLevelLimit - 1;
    }
    own[i] = max(0, kb - kt + 1);
    ktop[i] = kt;
}

// ---------------------------------------------------------------------------
// Exclusive scan of n ints (the cells each internal node holds -> the number
// of its first cell): tiles of 1024 scanned in place with their sums kept,
// the sums scanned by one block, then added back.
// ---------------------------------------------------------------------------
constexpr int kScanTile = 1024;
static_assert(kScanTile == kKeyScanTile, "the engine sizes the tile sums with kKeyScanTile");

__device__ __forceinline__ int block_excl_scan1024(int v, int *s, int *total)
{
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int st = 1; st < 1024; st <<= 1) {
        const int u = t >= st ? s[t - st] : 0;
        __syncthreads();
        s[t] += u;
        __syncthreads();
    }
    const int r = s[t] - v;
    if (total) *total = s[1023];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const int *__restrict__ in, int *__restrict__ out, int n,
                                                     int *__restrict__ sums)
{
    __shared__ int s[1024];
    const int i = blockIdx.x * kScanTile + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    int tot;
    const int e = block_excl_scan1024(v, s, &tot);
    if (i < n) out[i] = e;
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// one block: exclusive scan of the tile sums in place (each thread a run of them)
__global__ __launch_bounds__(1024) void k_scan_sums(int *__restrict__ sums, int ntiles)
{
    __shared__ int s[1024];
    const int t = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    int acc = 0;
    for (int k = 0; k < per; ++k) {
        const int j = t * per + k;
        if (j < ntiles) acc += sums[j];
    }
    int run = block_excl_scan1024(acc, s, nullptr);
    for (int k = 0; k < per; ++k) {
        const int j = t * per + k;
        if (j < ntiles) {
            const int c = sums[j];
            sums[j] = run;
            run += c;
        }
    }
}

__global__ __launch_bounds__(1024) void k_scan_add(int *__restrict__ out, int n, const int *__restrict__ sums)
{
    const int i = blockIdx.x * kScanTile + threadIdx.x;
    if (i < n) out[i] += sums[blockIdx.x];
}

// ---------------------------------------------------------------------------
// Cells and slots. Cell obase[i] + (k - ktop[i]) is internal node i's cell of
// level k; the root (level 0, held by binary node 0) is cell 0.
// ---------------------------------------------------------------------------
// The nearest binary ancestor (p itself included) that holds a cell. A node
// that holds none shares its digit level with its parent, so the walk is a
// few steps; it is bounded anyway, and -1 (an inconsistent tree) is flagged
// by the caller.
__device__ __forceinline__ int owning_ancestor(const int *__restrict__ iparent, const int *__restrict__ own, int p)
{
    for (int steps = 0; steps < 4 * kBuildLevelLimit * 3 && p >= 0; ++steps) {
        if (own[p] != 0) return p;
        p = iparent[p];
    }
    return -1;
}

struct KeyTree {
    const unsigned long long *key, *deep;
    const int *val, *iparent, *lparent, *dnode, *lo, *own, *ktop, *obase;
};

// the cell count, the pool check, and the root's record
__global__ void k_cell_count(Dev d, KeyTree kt)
{
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const int n = d.n;
    const int C = n >= 2 ? kt.obase[n - 2] + kt.own[n - 2] : 1;
    if (C > d.cap) {
        flag_error(ctl, kErrPool);
        return;
    }
    ctl->n_cells = C;
    d.cinfo[0] = make_int4(-1, 0, (int)ctl->epoch, 0);
}

__global__ __launch_bounds__(256) void k_cells(Dev d, KeyTree kt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = d.n;
    const Ctl *ctl = d.ctl;
    if (i >= n - 1 || ctl->error) return;
    const int own = kt.own[i];
    if (own == 0) return;
    const uint32_t ep = ctl->epoch;
    const int top = kt.ktop[i], base = kt.obase[i];
    for (int k = top; k < top + own; ++k) {
        const int cid = base + (k - top);
        if (k == 0) continue; // the root has no parent
        int par;
        if (k > top) {
            par = cid - 1;
        } else {
            const int y = owning_ancestor(kt.iparent, kt.own, kt.iparent[i]);
            if (y < 0) {
                flag_error(d.ctl, kErrSync);
                return;
            }
            par = kt.obase[y] + kt.own[y] - 1;
        }
        const int dg = key_digit(kt.key, kt.deep, n, kt.lo[i], k - 1);
        d.slot[(size_t)par * 8 + dg] = slot_tag(ep, n + cid);
        d.cinfo[cid] = make_int4(par, k | (dg << 8), (int)ep, 0);
    }
}

// Sorted body g into the deepest cell of its nearest holding ancestor; the
// running maximum depth (buildtree.cl:199), one atomic per block. The grid
// strides over the bodies: one atomic per wave on that single word cost 0.75
// ms at 4M bodies (65,536 waves).
constexpr int kLeafBlocks = 1024;

__global__ __launch_bounds__(256) void k_leaves(Dev d, KeyTree kt)
{
    __shared__ int part[4];
    const int n = d.n;
    Ctl *ctl = d.ctl;
    int depth = 0;
    if (!ctl->error) {
        const uint32_t ep = ctl->epoch;
        for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
            int par = 0, lv = 0, y = 0;
            if (n >= 2) {
                y = owning_ancestor(kt.iparent, kt.own, kt.lparent[g]);
                if (y >= 0) {
                    par = kt.obase[y] + kt.own[y] - 1;
                    lv = kt.ktop[y] + kt.own[y] - 1;
                }
            }
            if (y < 0) {
                flag_error(ctl, kErrSync);
                continue;
            }
            const int dg = key_digit(kt.key, kt.deep, n, g, lv);
            d.slot[(size_t)par * 8 + dg] = slot_tag(ep, kt.val[g]);
            depth = max(depth, lv + 1);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) depth = max(depth, __shfl_xor(depth, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = depth;
    __syncthreads();
    if (threadIdx.x == 0) {
        depth = max(max(part[0], part[1]), max(part[2], part[3]));
        if (depth > 0) __hip_atomic_fetch_max(&ctl->max_depth, depth, __ATOMIC_RELAXED, AGENT);
    }
}

// Stage 4 (sort.cl:12-76) on the key-sorted tree: the octant-path order is the
// build's sorted body order, so sorted[] is a copy of it and bnode[] gathers
// (x, y, z, m) in that order.
__global__ __launch_bounds__(256) void k_sort_keys(Dev d, const int *__restrict__ val)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= d.n || d.ctl->error) return;
    const int b = val[g];
    d.sorted[g] = b;
    d.bnode[g] = d.pm4[b];
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
hipError_t launch_build_keys(const Dev &d, hipStream_t st)
{
    const int n = d.n;
    const dim3 b256(256), g256((n + 255) / 256);
    const int ntiles = (n + kRsTile - 1) / kRsTile;
    unsigned long long *ka = d.kb_key[0], *kb = d.kb_key[1];
    int *va = d.kb_val[0], *vb = d.kb_val[1];
    hipLaunchKernelGGL(k_keys, g256, b256, 0, st, d, ka, va);
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
    // eight passes: the sorted keys are back in kb_key[0] / kb_val[0]; the
    // other pair and the binary tree's arrays (written later) serve long runs
    hipLaunchKernelGGL(k_deep_runs, g256, b256, 0, st, d, ka, va, d.kb_deep, d.kb_ipar, d.kb_lpar, kb, vb);
    KeyTree kt{ka, d.kb_deep, va, d.kb_ipar, d.kb_lpar, d.kb_dnode, d.kb_lo, d.kb_own, d.kb_ktop, d.kb_obase};
    if (n >= 2) {
        const dim3 gi((n - 1 + 255) / 256);
        hipLaunchKernelGGL(k_bintree, gi, b256, 0, st, d, ka, d.kb_deep, d.kb_ipar, d.kb_lpar, d.kb_dnode, d.kb_lo);
        hipLaunchKernelGGL(k_own, gi, b256, 0, st, d, d.kb_ipar, d.kb_dnode, d.kb_own, d.kb_ktop);
        const int m = n - 1, nt = (m + kScanTile - 1) / kScanTile;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanTile), 0, st, d.kb_own, d.kb_obase, m, d.kb_scan);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, d.kb_scan, nt);
        hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kScanTile), 0, st, d.kb_obase, m, d.kb_scan);
    }
    hipLaunchKernelGGL(k_cell_count, dim3(1), dim3(1), 0, st, d, kt);
    if (n >= 2) hipLaunchKernelGGL(k_cells, dim3((n - 1 + 255) / 256), b256, 0, st, d, kt);
    hipLaunchKernelGGL(k_leaves, dim3(min((n + 255) / 256, kLeafBlocks)), b256, 0, st, d, kt);
    return hipGetLastError();
}

hipError_t launch_sort_keys(const Dev &d, hipStream_t st)
{
    hipLaunchKernelGGL(k_sort_keys, dim3((d.n + 255) / 256), dim3(256), 0, st, d, d.kb_val[0]);
    return hipGetLastError();
}

} // namespace nb
