// nb_force.hip -- force traversal (calculateforce.cl:25-188) for gfx950.
//
// A block of 256 threads covers 256 consecutive sorted positions. A vote group
// is G consecutive lanes (G = 16, 32 or 64; 16 is the reference's work-group
// vote of calculateforce.cl:145 and 64 is one wave). Every group owns one
// stack in LDS; entries are (cell << 6 | depth). Popping a cell tests all of
// its (at most 8, compacted) children: bodies always interact, a child cell is
// accepted only if every lane of the group has d^2 + eps^2 >= dq[depth], and an
// opened cell is pushed. Which nodes interact therefore matches the reference
// exactly; only the order of the terms differs.
//
// Accumulation: the terms of one popped cell (<= 8) are summed in FP32 and the
// batch sum is added into a Kahan-compensated FP32 accumulator.
#include "nb_internal.h"

namespace nb {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

constexpr int kStack = 512; // >= 7 * 64 + 8 entries: the deepest possible push pattern

__device__ __forceinline__ void kahan_add(float &sum, float &comp, float x)
{
    const float y = x - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

template <int G>
__global__ __launch_bounds__(256) void k_force(Dev d, int fused)
{
    constexpr int kGroups = 256 / G;
    __shared__ unsigned stk[kGroups][kStack];
    __shared__ float dq_s[kMaxDepth];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int gl = tid / G;
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    if (ctl->max_depth > kMaxDepth) { // calculateforce.cl:69-73
        if (tid == 0) {
            int expected = 0;
            __hip_atomic_compare_exchange_strong(&ctl->error, &expected, (int)kErrDepth,
                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED, AGENT);
        }
        return;
    }
    if (tid < kMaxDepth) dq_s[tid] = ctl->dq[tid];
    __syncthreads();

    const int pos = blockIdx.x * 256 + tid;
    const bool valid = pos < n;
    const float4 me = valid ? d.nodes4[pos] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float eps2 = d.eps2;
    const unsigned long long gmask =
        (G == 64) ? ~0ull : (((1ull << (G & 63)) - 1ull) << ((lane / G) * G));
    const bool leader = (lane % G) == 0;

    if (leader) stk[gl][0] = 0; // root cell 0 at depth 0
    int sp = 1;
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    int ninter = 0;
    bool overflow = false;
    while (sp > 0) {
        __builtin_amdgcn_wave_barrier();
        const unsigned e = stk[gl][--sp];
        const int c = (int)(e >> 6);
        const int dep = (int)(e & 63u);
        const float thr = dq_s[dep];
        const int4 c0 = *reinterpret_cast<const int4 *>(d.cchild + (size_t)c * 8);
        const int4 c1 = *reinterpret_cast<const int4 *>(d.cchild + (size_t)c * 8 + 4);
        const int ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        float px = 0.f, py = 0.f, pz = 0.f;
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) break;
            const float4 q = d.nodes4[v]; // body in sorted order, or cell v - n
            const float dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            bool acc = v < n;
            if (!acc) {
                const bool pred = !valid || d2 >= thr;
                acc = (__ballot(pred) & gmask) == gmask;
            }
            if (acc) {
                const float ri = __builtin_amdgcn_rsqf(d2);
                const float f = q.w * (ri * ri * ri);
                px = fmaf(dx, f, px);
                py = fmaf(dy, f, py);
                pz = fmaf(dz, f, pz);
                ++cnt;
            } else {
                // the tree is at most kMaxDepth deep: anything else is corrupt
                if (sp >= kStack || v - n >= d.cap || dep + 1 >= kMaxDepth) {
                    overflow = true;
                    break;
                }
                if (leader) stk[gl][sp] = ((unsigned)(v - n) << 6) | (unsigned)(dep + 1);
                ++sp;
            }
        }
        kahan_add(sx, kx, px);
        kahan_add(sy, ky, py);
        kahan_add(sz, kz, pz);
        ninter += cnt;
        if (overflow) break;
    }
    if (overflow && leader) {
        int expected = 0;
        __hip_atomic_compare_exchange_strong(&ctl->error, &expected, (int)kErrStack, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, AGENT);
    }

    if (valid) {
        const int b = d.sorted[pos];
        const float fx = sx - kx, fy = sy - ky, fz = sz - kz;
        const float dt = d.dt;
        const float4 v0 = d.vel4[b];
        float vx = v0.x, vy = v0.y, vz = v0.z;
        if (ctl->step > 0) { // calculateforce.cl:174-179
            const float4 a0 = d.acc4[b];
            vx = vx + ((fx - a0.x) * dt) * 0.5f;
            vy = vy + ((fy - a0.y) * dt) * 0.5f;
            vz = vz + ((fz - a0.z) * dt) * 0.5f;
        }
        d.acc4[b] = make_float4(fx, fy, fz, 0.0f); // calculateforce.cl:183-185
        if (fused) { // integrate.cl:29-43 for this body
            const float dvx = (fx * dt) * 0.5f, dvy = (fy * dt) * 0.5f, dvz = (fz * dt) * 0.5f;
            const float vhx = vx + dvx, vhy = vy + dvy, vhz = vz + dvz;
            float4 p = me; // this body's (x, y, z, m), read from the sorted copy
            if (d.fp64) {
                const double4 q = d.pos64[b];
                const double x = fma((double)vhx, (double)dt, q.x);
                const double y = fma((double)vhy, (double)dt, q.y);
                const double z = fma((double)vhz, (double)dt, q.z);
                d.pos64[b] = make_double4(x, y, z, 0.0);
                p.x = (float)x; p.y = (float)y; p.z = (float)z;
            } else {
                p.x = fmaf(vhx, dt, p.x);
                p.y = fmaf(vhy, dt, p.y);
                p.z = fmaf(vhz, dt, p.z);
            }
            d.pm4[b] = p;
            vx = vhx + dvx;
            vy = vhy + dvy;
            vz = vhz + dvz;
        }
        d.vel4[b] = make_float4(vx, vy, vz, 0.0f);
    } else {
        ninter = 0;
    }
    // device interaction counter: one atomic per wave
    unsigned long long tot = (unsigned long long)ninter;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if (lane == 0 && tot) __hip_atomic_fetch_add(&ctl->interactions, tot, __ATOMIC_RELAXED, AGENT);
}

hipError_t launch_force(const Dev &d, int vote, int fused, hipStream_t st)
{
    const dim3 grid((d.n + 255) / 256), block(256);
    switch (vote) {
    case 16: hipLaunchKernelGGL(k_force<16>, grid, block, 0, st, d, fused); break;
    case 32: hipLaunchKernelGGL(k_force<32>, grid, block, 0, st, d, fused); break;
    case 64: hipLaunchKernelGGL(k_force<64>, grid, block, 0, st, d, fused); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace nb
