// nb_tree.hip -- bounding box, summarize, integrate and the vertex export,
// for gfx950 (the build and the sort are in nb_keys.hip).
//
// No kernel here waits on another workgroup. Summarize runs one launch per
// tree level, so every value it reads was written by an earlier launch: no
// fences, no waiting on other threads; the deepest levels, which hold few
// cells, share one single-block launch with a block barrier between levels.
// (On MI355X each XCD has its own L2; an agent-scope acquire/release costs an
// L2 invalidate / write-back, which is why the per-cell hand-offs of the
// reference's summarize are replaced by level launches.)
#include "nb_internal.h"

namespace nb {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

__device__ __forceinline__ int dec_slot(unsigned long long w, uint32_t ep)
{
    return ((uint32_t)(w >> 32) == ep) ? (int)(uint32_t)w : -1;
}

// Octant of a body relative to a cell centre: bit0 x, bit1 y, bit2 z, set when
// the centre is strictly below the body (buildtree.cl:66-68, 84-86).
__device__ __forceinline__ int octant(float cx, float cy, float cz, float bx, float by, float bz)
{
    return (cx < bx ? 1 : 0) | (cy < by ? 2 : 0) | (cz < bz ? 4 : 0);
}

__device__ __forceinline__ void set_error(Ctl *ctl, int code)
{
    int expected = 0;
    __hip_atomic_compare_exchange_strong(&ctl->error, &expected, code, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, AGENT);
}

// A cell is in the tree when its record carries this tree's epoch: the build
// writes cells 0 .. n_cells-1, and records past them keep an older epoch.
__device__ __forceinline__ bool cell_live(const Dev &d, int c, uint32_t ep, int *level)
{
    if (c == 0) {
        *level = 0;
        return true;
    }
    const int4 info = d.cinfo[c];
    *level = info.y & 255;
    return (uint32_t)info.z == ep && *level < kMaxDepth;
}

// ---------------------------------------------------------------------------
// Stage 1: bounding box (boundingbox.cl:20-199). Two launches: min/max per
// block (wave64 shuffle reduction, one partial per block), then one block
// combines the partials and sets up the root, the opening table and the
// per-step counters. The launch boundary orders the two, where the
// reference's last-block hand-off (boundingbox.cl:133-196) needs an
// agent-scope release per block: on MI355X an L2 write-back each, which cost
// about 40 us per step at 4M bodies against 11 us for the reads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__device__ void block_minmax(float v[6], float (*part)[6])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = wave_min(v[k]);
#pragma unroll
    for (int k = 3; k < 6; ++k) v[k] = wave_max(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) part[w][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < (int)(blockDim.x >> 6); ++j) {
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], part[j][k]);
#pragma unroll
            for (int k = 3; k < 6; ++k) v[k] = fmaxf(v[k], part[j][k]);
        }
    }
}

__device__ __forceinline__ void minmax_into(float v[6], float4 q)
{
    v[0] = fminf(v[0], q.x); v[1] = fminf(v[1], q.y); v[2] = fminf(v[2], q.z);
    v[3] = fmaxf(v[3], q.x); v[4] = fmaxf(v[4], q.y); v[5] = fmaxf(v[5], q.z);
}

__global__ __launch_bounds__(256) void k_bbox(Dev d)
{
    __shared__ float part[4][6];
    // a device error stops the pipeline until the host has seen it: the steps
    // that were already queued behind it change nothing
    if (d.ctl->error) return;
    const int n = d.n;
    // seeded with body 0 as the reference does (boundingbox.cl:44-58)
    const float4 b0 = d.pm4[0];
    float v[6] = {b0.x, b0.y, b0.z, b0.x, b0.y, b0.z};
    // Four independent loads in flight per lane, from one contiguous 4 KiB
    // span per wave and iteration: body j is read by wave (j / 256) mod W at
    // element j mod 256.
    const int lane = threadIdx.x & 63;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    long long i = (((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 256 + lane;
    for (; i + 192 < n; i += nwaves * 256) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = d.pm4[i + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u) minmax_into(v, q[u]);
    }
    // the span that crosses n (every later span of this wave starts past n)
    for (int u = 0; u < 4; ++u) {
        if (i + 64 * u >= n) break;
        minmax_into(v, d.pm4[i + 64 * u]);
    }
    block_minmax(v, part);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) d.bbox_part[6 * blockIdx.x + k] = v[k];
    }
}

// one block: the partials of k_bbox's `blocks` blocks
__global__ __launch_bounds__(256) void k_bbox_final(Dev d, int blocks)
{
    __shared__ float part[4][6];
    Ctl *c = d.ctl;
    if (c->error) return;
    float w[6] = {d.bbox_part[0], d.bbox_part[1], d.bbox_part[2], d.bbox_part[3], d.bbox_part[4],
                  d.bbox_part[5]};
    for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
        const float *q = d.bbox_part + 6 * b;
        w[0] = fminf(w[0], q[0]); w[1] = fminf(w[1], q[1]); w[2] = fminf(w[2], q[2]);
        w[3] = fmaxf(w[3], q[3]); w[4] = fmaxf(w[4], q[4]); w[5] = fmaxf(w[5], q[5]);
    }
    block_minmax(w, part);
    if (threadIdx.x == 0) {
        // root cell, boundingbox.cl:171-189
        c->root[0] = 0.5f * (w[0] + w[3]);
        c->root[1] = 0.5f * (w[1] + w[4]);
        c->root[2] = 0.5f * (w[2] + w[5]);
        const float radius = 0.5f * fmaxf(fmaxf(w[3] - w[0], w[4] - w[1]), w[5] - w[2]);
        c->radius = radius;
        // opening thresholds, calculateforce.cl:52-67 (FP32 quotient via a
        // double division rounded once = the correctly rounded FP32 quotient)
        float raw = radius * radius;
        if (d.theta2 > 0.0f) raw = (float)((double)raw / (double)d.theta2);
        for (int i = 0; i < kMaxDepth; ++i) {
            c->dq[i] = raw + d.eps2;
            raw = 0.25f * raw;
        }
        c->n_cells = 1; // the root is cell 0 (reference: _bottom = NUMBER_OF_NODES, :180)
        c->epoch += 1;  // every child slot of the last tree turns stale
        c->step += 1;   // boundingbox.cl:195
    }
}

// ---------------------------------------------------------------------------
// Live cells grouped by level: per-block histograms, one scan, a scatter.
// Blocks own contiguous ranges of cell numbers in both passes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_range(int C, int *lo, int *hi)
{
    const int per = (C + kLevelBlocks - 1) / kLevelBlocks;
    *lo = min(C, (int)blockIdx.x * per);
    *hi = min(C, *lo + per);
}

__global__ __launch_bounds__(256) void k_level_hist(Dev d)
{
    __shared__ int h[kMaxDepth];
    Ctl *ctl = d.ctl;
    if (threadIdx.x < kMaxDepth) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t ep = ctl->epoch;
    const int C = min(ctl->n_cells, d.cap);
    int lo, hi;
    block_range(C, &lo, &hi);
    if (!ctl->error) {
        for (int c = lo + threadIdx.x; c < hi; c += blockDim.x) {
            int lvl;
            const bool live = cell_live(d, c, ep, &lvl);
            if (live) atomicAdd(&h[lvl], 1);
            d.clvl[c] = live ? (unsigned char)lvl : kNotLive; // for the scatter: one coalesced byte
        }
    }
    __syncthreads();
    if (threadIdx.x < kMaxDepth) d.lvl_blockhist[blockIdx.x * kMaxDepth + threadIdx.x] = h[threadIdx.x];
}

// One block of 1024 threads turns the per-block level histograms into list
// offsets: thread (seg, L) owns level L in block segment seg (16 segments of
// kLevelBlocks / 16 blocks); segment sums, their prefix per level, the level
// starts, then each thread writes its segment's offsets.
constexpr int kScanSegs = 1024 / kMaxDepth;
static_assert(kLevelBlocks % kScanSegs == 0, "level blocks split evenly into scan segments");

__global__ __launch_bounds__(1024) void k_level_scan(Dev d)
{
    __shared__ int segsum[kScanSegs][kMaxDepth];
    __shared__ int base[kMaxDepth];
    const int L = threadIdx.x % kMaxDepth, seg = threadIdx.x / kMaxDepth;
    constexpr int per = kLevelBlocks / kScanSegs;
    int *h = d.lvl_blockhist + (size_t)seg * per * kMaxDepth + L;
    int sum = 0;
    for (int b = 0; b < per; ++b) sum += h[b * kMaxDepth];
    segsum[seg][L] = sum;
    __syncthreads();
    if (seg == 0) { // exclusive prefix over the segments of level L
        int acc = 0;
        for (int k = 0; k < kScanSegs; ++k) {
            const int v = segsum[k][L];
            segsum[k][L] = acc;
            acc += v;
        }
        base[L] = acc; // level total, for the moment
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Ctl *ctl = d.ctl;
        int acc = 0, levels = 0;
        for (int k = 0; k < kMaxDepth; ++k) {
            const int t = base[k];
            ctl->lvl_start[k] = acc;
            base[k] = acc;
            acc += t;
            if (t > 0) levels = k + 1;
        }
        ctl->lvl_start[kMaxDepth] = acc;
        ctl->levels = levels;
        ctl->live_cells = acc;
    }
    __syncthreads();
    int run = base[L] + segsum[seg][L];
    for (int b = 0; b < per; ++b) {
        const int v = h[b * kMaxDepth];
        h[b * kMaxDepth] = run;
        run += v;
    }
}

__global__ __launch_bounds__(256) void k_level_scatter(Dev d)
{
    __shared__ int cur[kMaxDepth];
    Ctl *ctl = d.ctl;
    if (threadIdx.x < kMaxDepth) cur[threadIdx.x] = d.lvl_blockhist[blockIdx.x * kMaxDepth + threadIdx.x];
    __syncthreads();
    if (ctl->error) return;
    const int C = min(ctl->n_cells, d.cap);
    int lo, hi;
    block_range(C, &lo, &hi);
    for (int c = lo + threadIdx.x; c < hi; c += blockDim.x) {
        const int lvl = d.clvl[c];
        if (lvl != kNotLive) d.lvl_list[atomicAdd(&cur[lvl], 1)] = c;
    }
}

// ---------------------------------------------------------------------------
// Stage 3: summarize (summarizetree.cl:15-179), one launch per level from the
// deepest up. Eight lanes per cell, one per child slot, so the children's
// records are fetched in parallel. A cell's centre of mass is summed in FP64
// from its children's results of the previous launch, over the eight slots
// (an empty slot adds 0) in the fixed pairwise order
// ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)), which the xor-shuffle
// reduction produces identically on every lane. Children are compacted to
// the front (summarizetree.cl:75-81) into the traversal list.
// ---------------------------------------------------------------------------
// Lane exchanges inside each group of 8 lanes as DPP moves (no LDS):
// quad_perm [1,0,3,2] = lane ^ 1, quad_perm [2,3,0,1] = lane ^ 2, and
// row_half_mirror (lane i <-> 7 - i) to join the two halves of 8. After the
// first two steps lanes 0-3 all hold (s0 + s1) + (s2 + s3) and lanes 4-7
// (s4 + s5) + (s6 + s7) (floating-point addition commutes exactly), so the
// third gives every lane the same pairwise sum as xor 4 would.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141;

template <int CTRL> __device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

template <int CTRL> __device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long u = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)(unsigned)u), hi = dpp_i<CTRL>((int)(unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double sum8(double v)
{
    v += dpp_d<kDppXor1>(v);
    v += dpp_d<kDppXor2>(v);
    v += dpp_d<kDppHalfMirror>(v);
    return v;
}

// Cell i of the level's list on the 8 lanes of one group, lane k = child slot
// k. Every lane of the group calls it (the exchanges need all eight); lanes of
// a group with i >= cells only take part in them.
template <bool QUAD>
__device__ __forceinline__ void summarize_cell(const Dev &d, int first, int i, int k, int cells, uint32_t ep)
{
    // products are rounded before they are summed, as in the oracle (no FMA
    // contraction: it would also fold a product into the first exchange step)
#pragma clang fp contract(off)
    const bool live = i < cells;
    const int n = d.n;
    int c = 0, v = -1;
    if (live) {
        c = d.lvl_list[first + i];
        v = dec_slot(d.slot[(size_t)c * 8 + k], ep);
    }
    double m = 0.0, x = 0.0, y = 0.0, z = 0.0;
    int cnt = 0;
    float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
    double q[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; // the child's own second moments
    double bc = 0.0;                                     // and its extent
    if (v >= 0 && v < n) {
        rec = d.pm4[v];
        m = rec.w;
        x = rec.x;
        y = rec.y;
        z = rec.z;
        cnt = 1;
    } else if (v >= n) {
        const double4 cc = d.com64[v - n];
        rec = make_float4((float)cc.x, (float)cc.y, (float)cc.z, (float)cc.w);
        m = cc.w;
        x = cc.x;
        y = cc.y;
        z = cc.z;
        cnt = d.count[v - n];
        if (QUAD) {
            const double4 *cq = reinterpret_cast<const double4 *>(d.cq64 + (size_t)(v - n) * 8);
            const double4 q0 = cq[0], q1 = cq[1];
            q[0] = q0.x; q[1] = q0.y; q[2] = q0.z; q[3] = q0.w;
            q[4] = q1.x; q[5] = q1.y; q[6] = q1.z;
            bc = q1.w;
        }
    }
    // compaction: position of this child among the cell's non-empty slots;
    // empty slots fill the tail
    const int gl = (int)(threadIdx.x & 63) & ~7;
    const unsigned gm = (unsigned)((__ballot(v >= 0) >> gl) & 0xffull);
    const int used = __popc(gm), below = __popc(gm & ((1u << k) - 1u));
    const int at = v >= 0 ? below : used + (k - below);
    const double sm = sum8(m), sx = sum8(m * x), sy = sum8(m * y), sz = sum8(m * z);
    int scnt = cnt;
    scnt += dpp_i<kDppXor1>(scnt);
    scnt += dpp_i<kDppXor2>(scnt);
    scnt += dpp_i<kDppHalfMirror>(scnt);
    // a cell of zero mass gets its centre at the origin instead of the
    // reference's 0 * inf = NaN (summarizetree.cl:161-167)
    const double inv = sm > 0.0 ? 1.0 / sm : 0.0;
    const double4 com = make_double4(sx * inv, sy * inv, sz * inv, sm);
    // the traversals and the sort read a slot's record only up to the first
    // empty slot: the empty ones keep stale records (a cell has 3 children on
    // average, so this skips well over half of the record writes)
    if (live) {
        d.cchild[(size_t)c * 8 + at] = v;
        if (v >= 0) d.cdata[(size_t)c * 8 + at] = rec;
        if (k == 0) {
            d.com64[c] = com;
            d.count[c] = scnt;
        }
    }
    if (QUAD) {
        // Second moments about the centre of mass for the softened quadrupole
        // of the learned mode: the traceless tensor Q = sum m (3 s s^T - |s|^2 I)
        // and the trace T = sum m |s|^2 (s = offset from the centre of mass):
        // per slot, the child's own moments plus, by the parallel-axis rule,
        // its mass at its offset from this cell's centre of mass, then summed
        // over the slots in the same pairwise order. Each child's moments also
        // sit next to its (x, y, z, m) for the traversal, as
        // (Qxx, Qyy, Qzz, T eps^2) and (Qxy, Qxz, Qyz, 0).
        // Extent b: the largest (|child centre - centre| + child extent) over
        // the slots, an upper bound of every body's distance from the centre
        // of mass (a body's extent is 0); each child's own b enters the
        // coarse traversal's tests as the threshold b^2 / theta_c^2 + eps^2
        // (an FP32 fmaf of b^2 rounded from FP64) and (m b^3)^(2/5), the
        // cell's factor of the relative test's threshold on s^2 (FP64
        // exp2(0.4 log2(m b^3)), rounded).
        // Explicit fma and no add after a product: the opening decisions that
        // follow from these values are bit-exact against the oracle.
        const double dx = x - com.x, dy = y - com.y, dz = z - com.z;
        double br = v >= 0 ? sqrt(fma(dx, dx, fma(dy, dy, dz * dz))) + bc : 0.0;
        br = fmax(br, dpp_d<kDppXor1>(br));
        br = fmax(br, dpp_d<kDppXor2>(br));
        br = fmax(br, dpp_d<kDppHalfMirror>(br));
        const double r2 = dx * dx + dy * dy + dz * dz;
        double t7[7];
        t7[0] = q[0] + m * (3.0 * dx * dx - r2);
        t7[1] = q[1] + m * (3.0 * dy * dy - r2);
        t7[2] = q[2] + m * (3.0 * dz * dz - r2);
        t7[3] = q[3] + m * 3.0 * dx * dy;
        t7[4] = q[4] + m * 3.0 * dx * dz;
        t7[5] = q[5] + m * 3.0 * dy * dz;
        t7[6] = q[6] + m * r2;
#pragma unroll
        for (int j = 0; j < 7; ++j) t7[j] = sum8(t7[j]);
        if (live) {
            // read for cell children only (a body meets no opening test and
            // has no second moments)
            if (v >= n) {
                const double eps2 = (double)d.eps2;
                d.cqa[(size_t)c * 8 + at] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)(q[6] * eps2));
                d.cqb[(size_t)c * 8 + at] = make_float4((float)q[3], (float)q[4], (float)q[5], 0.f);
                const double mb3 = m * ((bc * bc) * bc);
                d.cmac[(size_t)c * 8 + at] =
                    make_float2(fmaf((float)(bc * bc), d.it2_coarse, d.eps2), (float)exp2(0.4 * log2(mb3)));
            }
            if (k == 0) {
                double4 *cq = reinterpret_cast<double4 *>(d.cq64 + (size_t)c * 8);
                cq[0] = make_double4(t7[0], t7[1], t7[2], t7[3]);
                cq[1] = make_double4(t7[4], t7[5], t7[6], br);
            }
        }
    }
}

// One level; the cell count is read on the device (a level at or below the
// tree's depth has none), and the grid strides over the cells, so any grid
// size is correct.
template <bool QUAD>
__global__ __launch_bounds__(256) void k_summarize_level(Dev d, int level)
{
    const Ctl *ctl = d.ctl;
    if (level >= ctl->levels) return;
    const int first = ctl->lvl_start[level], cells = ctl->lvl_start[level + 1] - first;
    const uint32_t ep = ctl->epoch;
    const int k = threadIdx.x & 7;
    for (int i0 = blockIdx.x * (blockDim.x >> 3); i0 < cells; i0 += gridDim.x * (blockDim.x >> 3))
        summarize_cell<QUAD>(d, first, i0 + (threadIdx.x >> 3), k, cells, ep);
}

// Levels lo .. hi-1 (those the tree has), bottom-up, in one block: a barrier
// between levels makes each level's results visible to the next. With hi past
// the tree's depth it takes every level from lo down, however deep the tree.
template <bool QUAD>
__global__ __launch_bounds__(kTailThreads) void k_summarize_block(Dev d, int lo, int hi)
{
    const Ctl *ctl = d.ctl;
    const uint32_t ep = ctl->epoch;
    const int k = threadIdx.x & 7;
    for (int L = min(hi, ctl->levels) - 1; L >= lo; --L) {
        const int first = ctl->lvl_start[L], cells = ctl->lvl_start[L + 1] - first;
        for (int i0 = 0; i0 < cells; i0 += kTailThreads / 8)
            summarize_cell<QUAD>(d, first, i0 + (threadIdx.x >> 3), k, cells, ep);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stage 6: integrate (integrate.cl:17-70): kick with the acceleration just
// computed, drift, kick. FP64 positions when NB_FP64_POSITIONS is set.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_integrate(Dev d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n || d.ctl->error) return;
    const float dt = d.dt;
    const float4 a = d.acc4[i], v = d.vel4[i];
    float4 p = d.pm4[i];
    const float dvx = (a.x * dt) * 0.5f, dvy = (a.y * dt) * 0.5f, dvz = (a.z * dt) * 0.5f;
    const float vhx = v.x + dvx, vhy = v.y + dvy, vhz = v.z + dvz;
    if (d.fp64) {
        const double4 q = d.pos64[i];
        const double x = fma((double)vhx, (double)dt, q.x);
        const double y = fma((double)vhy, (double)dt, q.y);
        const double z = fma((double)vhz, (double)dt, q.z);
        d.pos64[i] = make_double4(x, y, z, 0.0);
        p.x = (float)x; p.y = (float)y; p.z = (float)z;
    } else {
        p.x = fmaf(vhx, dt, p.x);
        p.y = fmaf(vhy, dt, p.y);
        p.z = fmaf(vhz, dt, p.z);
    }
    d.pm4[i] = p;
    d.vel4[i] = make_float4(vhx + dvx, vhy + dvy, vhz + dvz, 0.0f);
}

// copyVertices (copyvertices.cl:7-17)
// Bodies first .. first + count - 1; the outputs are indexed from `first`
// (pos[i - first]), so a range can go into a staging buffer of its own size.
__global__ __launch_bounds__(256) void k_copy_vertices(Dev d, int first, int count, float4 *pos, float4 *vel)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const int i = first + j;
    if (pos) {
        const float4 p = d.pm4[i];
        pos[j] = make_float4(p.x, p.y, p.z, 1.0f);
    }
    if (vel) {
        const float4 v = d.vel4[i];
        vel[j] = make_float4(v.x, v.y, v.z, 1.0f);
    }
}

// Leaf membership: descend the child slots along each body's own octants
// until the slot that holds the body; record the octants. depth = -1 if the
// descent does not end at the body.
__global__ __launch_bounds__(256) void k_paths(Dev d, unsigned long long *paths3, int *depth)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.n) return;
    const int n = d.n;
    const uint32_t ep = d.ctl->epoch;
    const float4 me = d.pm4[b];
    float cx = d.ctl->root[0], cy = d.ctl->root[1], cz = d.ctl->root[2], r = d.ctl->radius;
    unsigned long long w[3] = {0, 0, 0};
    int p = 0, level = 0, result = -1;
    while (level < 63) {
        const int oct = octant(cx, cy, cz, me.x, me.y, me.z);
        w[level / 21] |= (unsigned long long)oct << (3 * (level % 21));
        const int v = dec_slot(d.slot[(size_t)p * 8 + oct], ep);
        if (v == b) { result = level + 1; break; }
        if (v < n || v - n >= d.cap) break;
        const float rh = r * 0.5f;
        cx = (cx - rh) + ((oct & 1) ? r : 0.0f);
        cy = (cy - rh) + ((oct & 2) ? r : 0.0f);
        cz = (cz - rh) + ((oct & 4) ? r : 0.0f);
        r = rh;
        p = v - n;
        ++level;
    }
    paths3[3 * (size_t)b + 0] = w[0];
    paths3[3 * (size_t)b + 1] = w[1];
    paths3[3 * (size_t)b + 2] = w[2];
    depth[b] = result;
}

// ---- launchers ---------------------------------------------------------------
hipError_t launch_bbox(const Dev &d, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_bbox, dim3(blocks), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, st, d, blocks);
    return hipGetLastError();
}
hipError_t launch_levels(const Dev &d, hipStream_t st)
{
    hipLaunchKernelGGL(k_level_hist, dim3(kLevelBlocks), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_level_scan, dim3(1), dim3(1024), 0, st, d);
    hipLaunchKernelGGL(k_level_scatter, dim3(kLevelBlocks), dim3(256), 0, st, d);
    return hipGetLastError();
}
// grid of a level launch: eight lanes per expected cell (at least one block)
static dim3 level_grid(int cells_hint) { return dim3((unsigned)(((cells_hint > 1 ? cells_hint : 1) * 8 + 255) / 256)); }

hipError_t launch_summarize_level(const Dev &d, int level, int cells_hint, hipStream_t st)
{
    if (d.quad)
        hipLaunchKernelGGL(k_summarize_level<true>, level_grid(cells_hint), dim3(256), 0, st, d, level);
    else
        hipLaunchKernelGGL(k_summarize_level<false>, level_grid(cells_hint), dim3(256), 0, st, d, level);
    return hipGetLastError();
}
hipError_t launch_summarize_block(const Dev &d, int lo, int hi, hipStream_t st)
{
    if (d.quad)
        hipLaunchKernelGGL(k_summarize_block<true>, dim3(1), dim3(kTailThreads), 0, st, d, lo, hi);
    else
        hipLaunchKernelGGL(k_summarize_block<false>, dim3(1), dim3(kTailThreads), 0, st, d, lo, hi);
    return hipGetLastError();
}
hipError_t launch_integrate(const Dev &d, hipStream_t st)
{
    hipLaunchKernelGGL(k_integrate, dim3((d.n + 255) / 256), dim3(256), 0, st, d);
    return hipGetLastError();
}
hipError_t launch_copy_vertices(const Dev &d, int first, int count, float4 *pos, float4 *vel, hipStream_t st)
{
    if (first < 0 || count < 1 || first > d.n - count) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy_vertices, dim3((count + 255) / 256), dim3(256), 0, st, d, first, count, pos, vel);
    return hipGetLastError();
}
hipError_t launch_paths(const Dev &d, unsigned long long *paths3, int *depth, hipStream_t st)
{
    hipLaunchKernelGGL(k_paths, dim3((d.n + 255) / 256), dim3(256), 0, st, d, paths3, depth);
    return hipGetLastError();
}

} // namespace nb
