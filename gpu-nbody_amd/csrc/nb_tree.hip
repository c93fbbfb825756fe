// nb_tree.hip -- bounding box, octree build, summarize, sort, integrate and
// the vertex export, for gfx950.
//
// Cross-workgroup communication inside a kernel uses HIP atomics at agent
// scope only: child slots are published with a release CAS and read with
// acquire loads; ready flags of summarize and sort are release stores read
// with acquire loads. MI355X keeps one L2 per XCD, so nothing here relies on
// plain stores becoming visible to another workgroup before a kernel ends.
#include "nb_internal.h"

namespace nb {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

__device__ __forceinline__ int dec_slot(unsigned long long w, uint32_t ep)
{
    return ((uint32_t)(w >> 32) == ep) ? (int)(uint32_t)w : -1;
}

__device__ __forceinline__ unsigned long long tag(uint32_t ep, int v)
{
    return ((unsigned long long)ep << 32) | (uint32_t)v;
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

// ---------------------------------------------------------------------------
// Stage 1: bounding box (boundingbox.cl:20-199). Grid-stride min/max, wave64
// shuffle reduction, one partial per block, and the last block to arrive
// combines the partials and sets up the root, the opening table and the
// per-step counters.
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

__global__ __launch_bounds__(256) void k_bbox(Dev d)
{
    __shared__ float part[4][6];
    __shared__ int is_last;
    const int n = d.n;
    // seeded with body 0 as the reference does (boundingbox.cl:44-58)
    float v[6] = {d.x[0], d.y[0], d.z[0], d.x[0], d.y[0], d.z[0]};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float px = d.x[i], py = d.y[i], pz = d.z[i];
        v[0] = fminf(v[0], px); v[1] = fminf(v[1], py); v[2] = fminf(v[2], pz);
        v[3] = fmaxf(v[3], px); v[4] = fmaxf(v[4], py); v[5] = fmaxf(v[5], pz);
    }
    block_minmax(v, part);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) d.bbox_part[6 * blockIdx.x + k] = v[k];
        const uint32_t prev = __hip_atomic_fetch_add(&d.ctl->bbox_count, 1u, __ATOMIC_ACQ_REL, AGENT);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    float w[6] = {d.bbox_part[0], d.bbox_part[1], d.bbox_part[2], d.bbox_part[3], d.bbox_part[4],
                  d.bbox_part[5]};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        const float *q = d.bbox_part + 6 * b;
        w[0] = fminf(w[0], q[0]); w[1] = fminf(w[1], q[1]); w[2] = fminf(w[2], q[2]);
        w[3] = fmaxf(w[3], q[3]); w[4] = fmaxf(w[4], q[4]); w[5] = fmaxf(w[5], q[5]);
    }
    __syncthreads();
    block_minmax(w, part);
    if (threadIdx.x == 0) {
        Ctl *c = d.ctl;
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
        c->live_cells = 0;
        c->epoch += 1;  // every child slot and ready flag of the last tree turns stale
        c->step += 1;   // boundingbox.cl:195
        c->bbox_count = 0;
    }
}

// ---------------------------------------------------------------------------
// Stage 2: octree build (buildtree.cl:12-200), lock-free.
//
// A child slot is a 64-bit word (epoch << 32 | child); a word of another
// epoch reads as empty, so the previous step's tree needs no clearing. Slot
// states only move forward within a step: empty -> body -> cell. A body goes
// into an empty slot with one CAS. A slot that holds a body d is split by
// taking a fresh cell c (index larger than the parent's), placing d in c's
// slot for its octant, and publishing c with a release CAS of d -> N+c; the
// inserting body then continues inside c. A lost race leaves c unpublished:
// it is cleared and kept as the thread's spare for a later split. Cell
// centres are never read from memory: each thread derives them from the
// root along its own path with the reference's arithmetic (buildtree.cl:124-136).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_build(Dev d)
{
    const int n = d.n, cap = d.cap;
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const uint32_t ep = ctl->epoch;
    const float radius = ctl->radius;
    const float rx = ctl->root[0], ry = ctl->root[1], rz = ctl->root[2];
    int my_max = 0;
    int spare = -1;
    bool failed = false;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < n && !failed;
         b += gridDim.x * blockDim.x) {
        const float bx = d.x[b], by = d.y[b], bz = d.z[b];
        int p = 0, level = 0;
        float cx = rx, cy = ry, cz = rz, r = radius;
        for (;;) {
            const int oct = octant(cx, cy, cz, bx, by, bz);
            unsigned long long *sp = d.slot + (size_t)p * 8 + oct;
            unsigned long long w = __hip_atomic_load(sp, __ATOMIC_ACQUIRE, AGENT);
            int v = dec_slot(w, ep);
            int next = -1; // cell to descend into
            float ncx = 0.f, ncy = 0.f, ncz = 0.f;
            const float rh = r * 0.5f;
            for (int tries = 0;; ++tries) {
                if (tries > 16) { set_error(ctl, kErrSync); failed = true; break; }
                if (v >= n) { next = v - n; break; }
                if (v < 0) {
                    if (__hip_atomic_compare_exchange_strong(sp, &w, tag(ep, b), __ATOMIC_ACQUIRE,
                                                             __ATOMIC_ACQUIRE, AGENT)) {
                        my_max = max(my_max, level + 1);
                        break;
                    }
                    v = dec_slot(w, ep);
                    continue;
                }
                // slot holds body v: split it
                if (level + 1 >= kBuildLevelLimit) { set_error(ctl, kErrDepth); failed = true; break; }
                int c;
                if (spare > p) {
                    c = spare;
                    spare = -1;
                } else {
                    c = __hip_atomic_fetch_add(&ctl->n_cells, 1, __ATOMIC_RELAXED, AGENT);
                    if (c >= cap) { set_error(ctl, kErrPool); failed = true; break; }
                    if (c <= p) { set_error(ctl, kErrSync); failed = true; break; }
                }
                ncx = (cx - rh) + ((oct & 1) ? r : 0.0f);
                ncy = (cy - rh) + ((oct & 2) ? r : 0.0f);
                ncz = (cz - rh) + ((oct & 4) ? r : 0.0f);
                const int od = octant(ncx, ncy, ncz, d.x[v], d.y[v], d.z[v]);
                unsigned long long *csl = d.slot + (size_t)c * 8 + od;
                __hip_atomic_store(csl, tag(ep, v), __ATOMIC_RELAXED, AGENT);
                if (__hip_atomic_compare_exchange_strong(sp, &w, tag(ep, n + c), __ATOMIC_ACQ_REL,
                                                         __ATOMIC_ACQUIRE, AGENT)) {
                    my_max = max(my_max, level + 2);
                    next = c;
                    break;
                }
                __hip_atomic_store(csl, tag(ep, (int)kEmpty), __ATOMIC_RELAXED, AGENT);
                spare = c;
                v = dec_slot(w, ep);
            }
            if (failed || next < 0) break;
            if (next <= p || next >= cap) { set_error(ctl, kErrSync); failed = true; break; }
            // descend (buildtree.cl:277-289): child centre from this cell's centre
            cx = (cx - rh) + ((oct & 1) ? r : 0.0f);
            cy = (cy - rh) + ((oct & 2) ? r : 0.0f);
            cz = (cz - rh) + ((oct & 4) ? r : 0.0f);
            r = rh;
            p = next;
            ++level;
            if (level >= kBuildLevelLimit) { set_error(ctl, kErrDepth); failed = true; break; }
        }
    }
    // running maximum depth (buildtree.cl:199)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) my_max = max(my_max, __shfl_xor(my_max, o));
    if ((threadIdx.x & 63) == 0 && my_max > 0)
        __hip_atomic_fetch_max(&ctl->max_depth, my_max, __ATOMIC_RELAXED, AGENT);
}

// ---------------------------------------------------------------------------
// Stage 3: summarize (summarizetree.cl:15-179). Persistent grid; each thread
// walks its cells from the highest index down. A cell is finished when every
// child cell carries this step's ready flag; its centre of mass is then summed
// in FP64 in child-slot order, stored, and its own flag is released. Children
// always have larger indices than their parent, so the highest unfinished cell
// can always proceed and the polling terminates (bounded by spin_limit).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_summarize(Dev d)
{
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const int n = d.n;
    const uint32_t ep = ctl->epoch;
    const int C = min(ctl->n_cells, d.cap);
    const int T = gridDim.x * blockDim.x;
    const int limit = ctl->spin_limit;
    int c = C - 1 - (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int spins = 0;
    int live = 0;
    while (c >= 0) {
        const unsigned long long *sl = d.slot + (size_t)c * 8;
        int ch[8];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) ch[k] = dec_slot(sl[k], ep);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (ch[k] >= n) {
                const uint32_t r = __hip_atomic_load(d.ready + (ch[k] - n), __ATOMIC_RELAXED, AGENT);
                ok = ok && (r == ep);
            }
        }
        if (!ok) {
            if (++spins > limit) {
                set_error(ctl, kErrSync);
                return;
            }
            if ((spins & 1023) == 0 && __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, AGENT)) return;
            continue;
        }
        // one acquire for all child flags read above (summarizetree.cl:86-102)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        spins = 0;
        double m = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
        int cnt = 0, used = 0;
        int *out = d.cchild + (size_t)c * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) continue;
            out[used++] = v; // children compacted to the front (summarizetree.cl:75-81)
            if (v < n) {
                const double mb = (double)d.m[v];
                m += mb;
                sx = fma(mb, (double)d.x[v], sx);
                sy = fma(mb, (double)d.y[v], sy);
                sz = fma(mb, (double)d.z[v], sz);
                cnt += 1;
            } else {
                const double4 q = d.com64[v - n];
                m += q.w;
                sx = fma(q.w, q.x, sx);
                sy = fma(q.w, q.y, sy);
                sz = fma(q.w, q.z, sz);
                cnt += d.count[v - n];
            }
        }
        for (int k = used; k < 8; ++k) out[k] = -1;
        // a cell of zero mass gets its centre at the origin instead of the
        // reference's 0 * inf = NaN (summarizetree.cl:161-167)
        const double inv = m > 0.0 ? 1.0 / m : 0.0;
        const double4 com = make_double4(sx * inv, sy * inv, sz * inv, m);
        d.com64[c] = com;
        d.cnode[c] = make_float4((float)com.x, (float)com.y, (float)com.z, (float)m);
        d.count[c] = cnt;
        __hip_atomic_store(d.ready + c, ep, __ATOMIC_RELEASE, AGENT); // summarizetree.cl:170-172
        live += cnt > 0 ? 1 : 0;
        c -= T;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) live += __shfl_xor(live, o);
    if ((threadIdx.x & 63) == 0 && live) __hip_atomic_fetch_add(&ctl->live_cells, live, __ATOMIC_RELAXED, AGENT);
}

// ---------------------------------------------------------------------------
// Stage 4: sort (sort.cl:12-76). Persistent grid; each thread walks its cells
// from the lowest index up. A cell whose start offset has arrived hands
// consecutive offsets to its children in octant order: child cells get a
// start and a ready flag, bodies get their sorted position, their (x,y,z,m)
// in sorted order, and the traversal child list is rewritten to reference that
// position. The result is the octant-path (Morton) order of the bodies.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sort(Dev d)
{
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    const int n = d.n;
    const uint32_t ep = ctl->epoch;
    const int C = min(ctl->n_cells, d.cap);
    const int T = gridDim.x * blockDim.x;
    const int limit = ctl->spin_limit;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int spins = 0;
    while (c < C) {
        if (d.count[c] == 0) { // allocated but never linked into the tree
            c += T;
            continue;
        }
        int st;
        if (c == 0) {
            st = 0; // root, boundingbox.cl:189
        } else {
            const uint32_t r = __hip_atomic_load(d.sready + c, __ATOMIC_RELAXED, AGENT); // sort.cl:215
            if (r != ep) {
                if (++spins > limit) {
                    set_error(ctl, kErrSync);
                    return;
                }
                if ((spins & 1023) == 0 && __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, AGENT)) return;
                continue;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            st = d.start[c];
        }
        spins = 0;
        int *cc = d.cchild + (size_t)c * 8;
        int ch[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) ch[k] = cc[k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) break;
            if (v >= n) { // sort.cl:223-237
                d.start[v - n] = st;
                st += d.count[v - n];
            } else { // sort.cl:238-245
                d.sorted[st] = v;
                d.bnode[st] = make_float4(d.x[v], d.y[v], d.z[v], d.m[v]);
                cc[k] = st;
                ++st;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) break;
            if (v >= n) __hip_atomic_store(d.sready + (v - n), ep, __ATOMIC_RELAXED, AGENT);
        }
        c += T;
    }
}

// ---------------------------------------------------------------------------
// Stage 6: integrate (integrate.cl:17-70): kick with the acceleration just
// computed, drift, kick. FP64 positions when NB_FP64_POSITIONS is set.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_integrate(Dev d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const float dt = d.dt;
    const float dvx = (d.ax[i] * dt) * 0.5f, dvy = (d.ay[i] * dt) * 0.5f, dvz = (d.az[i] * dt) * 0.5f;
    const float vhx = d.vx[i] + dvx, vhy = d.vy[i] + dvy, vhz = d.vz[i] + dvz;
    if (d.fp64) {
        const double x = fma((double)vhx, (double)dt, d.x64[i]);
        const double y = fma((double)vhy, (double)dt, d.y64[i]);
        const double z = fma((double)vhz, (double)dt, d.z64[i]);
        d.x64[i] = x; d.y64[i] = y; d.z64[i] = z;
        d.x[i] = (float)x; d.y[i] = (float)y; d.z[i] = (float)z;
    } else {
        d.x[i] = fmaf(vhx, dt, d.x[i]);
        d.y[i] = fmaf(vhy, dt, d.y[i]);
        d.z[i] = fmaf(vhz, dt, d.z[i]);
    }
    d.vx[i] = vhx + dvx;
    d.vy[i] = vhy + dvy;
    d.vz[i] = vhz + dvz;
}

// copyVertices (copyvertices.cl:7-17)
__global__ __launch_bounds__(256) void k_copy_vertices(Dev d, float4 *pos, float4 *vel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    if (pos) pos[i] = make_float4(d.x[i], d.y[i], d.z[i], 1.0f);
    if (vel) vel[i] = make_float4(d.vx[i], d.vy[i], d.vz[i], 1.0f);
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
    const float bx = d.x[b], by = d.y[b], bz = d.z[b];
    float cx = d.ctl->root[0], cy = d.ctl->root[1], cz = d.ctl->root[2], r = d.ctl->radius;
    unsigned long long w[3] = {0, 0, 0};
    int p = 0, level = 0, result = -1;
    while (level < 63) {
        const int oct = octant(cx, cy, cz, bx, by, bz);
        w[level / 21] |= (unsigned long long)oct << (3 * (level % 21));
        const int v = dec_slot(d.slot[(size_t)p * 8 + oct], ep);
        if (v == b) { result = level + 1; break; }
        if (v < n || v - n <= p || v - n >= d.cap) break;
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
    return hipGetLastError();
}
hipError_t launch_build(const Dev &d, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_build, dim3(blocks), dim3(256), 0, st, d);
    return hipGetLastError();
}
hipError_t launch_summarize(const Dev &d, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_summarize, dim3(blocks), dim3(256), 0, st, d);
    return hipGetLastError();
}
hipError_t launch_sort(const Dev &d, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_sort, dim3(blocks), dim3(256), 0, st, d);
    return hipGetLastError();
}
hipError_t launch_integrate(const Dev &d, hipStream_t st)
{
    hipLaunchKernelGGL(k_integrate, dim3((d.n + 255) / 256), dim3(256), 0, st, d);
    return hipGetLastError();
}
hipError_t launch_copy_vertices(const Dev &d, float4 *pos, float4 *vel, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_vertices, dim3((d.n + 255) / 256), dim3(256), 0, st, d, pos, vel);
    return hipGetLastError();
}
hipError_t launch_paths(const Dev &d, unsigned long long *paths3, int *depth, hipStream_t st)
{
    hipLaunchKernelGGL(k_paths, dim3((d.n + 255) / 256), dim3(256), 0, st, d, paths3, depth);
    return hipGetLastError();
}

static int blocks_per_cu(const void *kernel)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess) return 1;
    return nb;
}
int summarize_blocks_per_cu(int) { return blocks_per_cu((const void *)k_summarize); }
int sort_blocks_per_cu(int) { return blocks_per_cu((const void *)k_sort); }

} // namespace nb
