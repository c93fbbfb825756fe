// nb_force.hip -- force traversal (calculateforce.cl:25-188) for gfx950, and
// the FP64 direct sum used to measure force errors.
//
// A block is one wave: 64 consecutive sorted positions. A vote group
// is G consecutive lanes (G = 16, 32 or 64; 16 is the reference's work-group
// vote of calculateforce.cl:145 and 64 is one wave). Every group owns one
// stack in LDS; entries are (cell << 6 | depth). Popping a cell tests all of
// its (at most 8, compacted) children: bodies always interact, a child cell is
// accepted only if every lane of the group has d^2 + eps^2 >= dq[depth], and an
// opened cell is pushed. Which nodes interact therefore matches the reference
// exactly; only the order of the terms differs.
//
// A popped cell's children indices and their (x, y, z, m) are contiguous
// (cchild, cdata: 32 + 128 bytes). With G = 64 the whole traversal state is
// wave-uniform: the cell comes off the LDS stack into a scalar register and
// its 160 bytes arrive through scalar loads, so the vector ALU only does the
// per-body arithmetic and the opening vote is one ballot.
//
// Accumulation: the terms of one popped cell (<= 8) are summed in FP32 and the
// batch sum is added into a Kahan-compensated FP32 accumulator.
#include "nb_device.h"

namespace nb {


// ---------------------------------------------------------------------------
// Vote width 64: one wave, one group, uniform traversal. kForceWaves waves
// (groups) per block: they share nothing, and a block of one wave frees its
// slot as soon as its group is done, where a block of four holds its CU's
// share until its slowest group ends (8.97 -> 8.67 ms at 4M clustered bodies).
// ---------------------------------------------------------------------------
constexpr int kForceWaves = 1;

__global__ __launch_bounds__(64 * kForceWaves) void k_force64(Dev d, const int *__restrict__ cchild,
                                                 const float4 *__restrict__ cdata, int mode)
{
    __shared__ unsigned stk[kForceWaves][kStack];
    __shared__ float dq_s[kMaxDepth];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    if (ctl->max_depth > kMaxDepth) { // calculateforce.cl:69-73
        if (tid == 0) flag_error(ctl, kErrDepth);
        return;
    }
    for (int i = tid; i < kMaxDepth; i += 64 * kForceWaves) dq_s[i] = ctl->dq[i];
    const int step = ctl->step;
    __syncthreads();

    const int pos = blockIdx.x * (64 * kForceWaves) + tid;
    const bool valid = pos < n;
    // a lane past the last body sits at infinity: its d^2 is infinite, so it
    // accepts every cell and never holds up the vote (its sums are discarded)
    const float4 me = valid ? d.pm4[d.sorted[pos]] : make_float4(kFar, kFar, kFar, 0.f);
    const float eps2 = d.eps2;
    const int cap = d.cap;

    if (lane == 0) stk[wv][0] = 0; // root cell 0 at depth 0
    int sp = 1;
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    long long ninter = 0; // per wave: every lane meets the same nodes
    bool bad = false;
    while (sp > 0) {
        __builtin_amdgcn_wave_barrier();
        --sp;
        const unsigned e = __builtin_amdgcn_readfirstlane(stk[wv][sp]);
        const int c = (int)(e >> 6);
        const int dep = (int)(e & 63u);
        const float thr = dq_s[dep];
        const int *chp = cchild + (size_t)c * 8;
        const float4 *cdp = cdata + (size_t)c * 8;
        int ch[8];
        float4 q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ch[k] = __builtin_amdgcn_readfirstlane(chp[k]);
            q[k] = cdp[k];
        }
        float px = 0.f, py = 0.f, pz = 0.f;
        int cnt = 0;
        int npush = 0;
        bool badv = false;
        unsigned push = 0;
        // Every branch in this loop is wave-uniform and none leaves it early
        // (children are compacted to the front: past the first negative entry
        // every one is negative). An opened cell's stack entry goes into lane
        // `npush` of `push` by a lane write, and the opened children are stored
        // after the loop in one go: no lane-divergent store inside the loop,
        // which otherwise makes the compiler copy the running sums at every
        // child.
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) continue;
            const float dx = q[k].x - me.x, dy = q[k].y - me.y, dz = q[k].z - me.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            bool acc = v < n;
            if (!acc) acc = wave_all(d2 >= thr);
            if (acc) {
                const float ri = __builtin_amdgcn_rsqf(d2);
                const float f = q[k].w * (ri * ri * ri);
                px = fmaf(dx, f, px);
                py = fmaf(dy, f, py);
                pz = fmaf(dz, f, pz);
                ++cnt;
            } else {
                push = lane_write(push, ((unsigned)(v - n) << 6) | (unsigned)(dep + 1), npush);
                ++npush;
                badv |= v - n >= cap;
            }
        }
        // the opened children in child order (the last one is popped first);
        // the tree is at most kMaxDepth deep and the pool holds `cap` cells:
        // anything else is corrupt
        if (npush) {
            if (sp + npush > kStack || badv || dep + 1 >= kMaxDepth) {
                bad = true;
            } else {
                if (lane < npush) stk[wv][sp + lane] = push;
                sp += npush;
            }
        }
        kahan_add(sx, kx, px);
        kahan_add(sy, ky, py);
        kahan_add(sz, kz, pz);
        ninter += cnt;
        if (bad) break;
    }
    if (bad && lane == 0) flag_error(ctl, kErrStack);
    if (valid) finish_body(d, pos, me, sx - kx, sy - ky, sz - kz, mode, step);
    // device interaction counter: one atomic per wave
    const unsigned long long nvalid = (unsigned long long)__popcll(__ballot(valid));
    if (lane == 0 && ninter && nvalid)
        __hip_atomic_fetch_add(&ctl->interactions, (unsigned long long)ninter * nvalid, __ATOMIC_RELAXED, AGENT);
}

// ---------------------------------------------------------------------------
// Vote width 16 or 32: several groups per wave, each with its own stack.
// ---------------------------------------------------------------------------
// threads per block of the vote-16/32 traversal: one wave (its groups share
// nothing; blocks of 256 held a CU's share until their slowest group ended:
// vote 16 27.6 -> 25.6 ms, vote 32 18.6 -> 18.0 ms at 4M clustered bodies)
constexpr int kGThreads = 64;

template <int G>
__global__ __launch_bounds__(kGThreads) void k_force_g(Dev d, int mode)
{
    constexpr int kGroups = kGThreads / G;
    __shared__ unsigned stk[kGroups][kStack];
    __shared__ float dq_s[kMaxDepth];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int gl = tid / G;
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    if (ctl->max_depth > kMaxDepth) { // calculateforce.cl:69-73
        if (tid == 0) flag_error(ctl, kErrDepth);
        return;
    }
    for (int i = tid; i < kMaxDepth; i += kGThreads) dq_s[i] = ctl->dq[i];
    const int step = ctl->step;
    __syncthreads();

    const int pos = blockIdx.x * kGThreads + tid;
    const bool valid = pos < n;
    // a lane past the last body sits at infinity: its d^2 is infinite, so it
    // accepts every cell and never holds up the vote (its sums are discarded)
    const float4 me = valid ? d.pm4[d.sorted[pos]] : make_float4(kFar, kFar, kFar, 0.f);
    const float eps2 = d.eps2;
    const unsigned long long gmask = ((1ull << G) - 1ull) << ((lane / G) * G);
    const int glane = lane % G; // this lane's place in its group
    const bool leader = glane == 0;

    if (leader) stk[gl][0] = 0; // root cell 0 at depth 0
    int sp = 1;
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    int ninter = 0;
    bool bad = false;
    while (sp > 0) {
        __builtin_amdgcn_wave_barrier();
        const unsigned e = stk[gl][--sp];
        const int c = (int)(e >> 6);
        const int dep = (int)(e & 63u);
        const float thr = dq_s[dep];
        const int4 c0 = *reinterpret_cast<const int4 *>(d.cchild + (size_t)c * 8);
        const int4 c1 = *reinterpret_cast<const int4 *>(d.cchild + (size_t)c * 8 + 4);
        const int ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        const float4 *cdp = d.cdata + (size_t)c * 8;
        float px = 0.f, py = 0.f, pz = 0.f;
        int cnt = 0;
        int npush = 0;
        bool badv = false;
        unsigned push = 0;
        // as in k_force64: no early exit, and the opened children's stack
        // entries gathered in lanes 0 .. npush-1 of the group by a select and
        // stored after the loop
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) continue;
            const float4 q = cdp[k];
            const float dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            bool acc = v < n;
            if (!acc) acc = (__builtin_amdgcn_ballot_w64(d2 >= thr) & gmask) == gmask;
            if (acc) {
                const float ri = __builtin_amdgcn_rsqf(d2);
                const float f = q.w * (ri * ri * ri);
                px = fmaf(dx, f, px);
                py = fmaf(dy, f, py);
                pz = fmaf(dz, f, pz);
                ++cnt;
            } else {
                push = glane == npush ? ((unsigned)(v - n) << 6) | (unsigned)(dep + 1) : push;
                ++npush;
                badv |= v - n >= d.cap;
            }
        }
        if (npush) {
            if (sp + npush > kStack || badv || dep + 1 >= kMaxDepth) {
                bad = true;
            } else {
                if (glane < npush) stk[gl][sp + glane] = push;
                sp += npush;
            }
        }
        kahan_add(sx, kx, px);
        kahan_add(sy, ky, py);
        kahan_add(sz, kz, pz);
        ninter += cnt;
        if (bad) break;
    }
    if (bad && leader) flag_error(ctl, kErrStack);
    if (valid) finish_body(d, pos, me, sx - kx, sy - ky, sz - kz, mode, step);
    unsigned long long tot = valid ? (unsigned long long)ninter : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
    if (lane == 0 && tot) __hip_atomic_fetch_add(&ctl->interactions, tot, __ATOMIC_RELAXED, AGENT);
}

hipError_t launch_force(const Dev &d, int vote, int mode, hipStream_t st)
{
    const dim3 grid((d.n + kGThreads - 1) / kGThreads), block(kGThreads);
    switch (vote) {
    case 16: hipLaunchKernelGGL(k_force_g<16>, grid, block, 0, st, d, mode); break;
    case 32: hipLaunchKernelGGL(k_force_g<32>, grid, block, 0, st, d, mode); break;
    case 64:
        hipLaunchKernelGGL(k_force64, dim3((d.n + 64 * kForceWaves - 1) / (64 * kForceWaves)), dim3(64 * kForceWaves),
                           0, st, d, (const int *)d.cchild, (const float4 *)d.cdata, mode);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 direct sum on sampled bodies (the error yardstick; G = 1), with the
// softened potential sum_{j != i} m_j / sqrt(r^2 + eps^2) in the w component
// when PHI (else w = 0). Block (t, s) sums the bodies [s * chunk, (s + 1) *
// chunk) for 64 sampled targets: 4 threads per target stride through the
// chunk, whose bodies are staged through LDS (converted to FP64) once for all
// 64 targets. 1 / sqrt(r^2) starts from the FP64 hardware estimate
// (v_rsq_f64) and takes one FP64 Newton step, y (3/2 - r^2 y^2 / 2), which
// squares its relative error (the parity test holds it to 1e-12). Partial
// sums per (target, chunk) are combined by a second kernel in a fixed order,
// so the result is deterministic.
// ---------------------------------------------------------------------------
constexpr int kDirChunks = 64;  // source chunks
constexpr int kDirTargets = 64; // targets per block
constexpr int kDirTile = 256;   // bodies staged per LDS tile

template <bool PHI>
__global__ __launch_bounds__(256) void k_direct64(Dev d, const int *__restrict__ idx, int ns, double4 *part)
{
    __shared__ double4 tile[kDirTile];
    __shared__ double4 red[256];
    const int n = d.n;
    const int t0 = blockIdx.x * kDirTargets;
    const int chunk = (n + kDirChunks - 1) / kDirChunks;
    const int lo = min(n, (int)blockIdx.y * chunk), hi = min(n, lo + chunk);
    const int tl = threadIdx.x & (kDirTargets - 1); // target of this thread
    // 0..3: which quarter of each tile (the wave's index: uniform, so the
    // source loop below is counted in scalar registers)
    const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kDirTargets);
    const int tgt = t0 + tl;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    int self = -1;
    if (tgt < ns) {
        const int b = idx[tgt];
        if (b >= 0 && b < n) {
            const float4 p = d.pm4[b];
            tx = p.x; ty = p.y; tz = p.z;
            self = b;
        }
    }
    const double eps2 = (double)d.eps2;
    double ax = 0.0, ay = 0.0, az = 0.0, phi = 0.0;
    for (int base = lo; base < hi; base += kDirTile) {
        __syncthreads();
        const int j = base + threadIdx.x;
        const float4 pj = j < hi ? d.pm4[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        tile[threadIdx.x] = make_double4(pj.x, pj.y, pj.z, pj.w);
        __syncthreads();
        const int cnt = min(kDirTile, hi - base);
        // one source's terms, added in source order
        auto pair = [&](int k) {
            const double4 s = tile[k];
            const double dx = s.x - tx, dy = s.y - ty, dz = s.z - tz;
            const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
            double ri = __builtin_amdgcn_rsq(r2);
            ri = ri * fma(-0.5 * r2, ri * ri, 1.5);
            const double f = s.w * ri * ri * ri;
            ax = fma(dx, f, ax);
            ay = fma(dy, f, ay);
            az = fma(dz, f, az);
            if (PHI) phi = fma(s.w, base + k != self ? ri : 0.0, phi);
        };
        // four sources per round: their loads and arithmetic overlap, only the
        // final adds are chained (the same order as one at a time)
        int k = sub;
        for (; k + 12 < cnt; k += 16) {
            pair(k);
            pair(k + 4);
            pair(k + 8);
            pair(k + 12);
        }
        for (; k < cnt; k += 4) pair(k);
    }
    red[threadIdx.x] = make_double4(ax, ay, az, phi);
    __syncthreads();
    if (sub == 0 && tgt < ns) {
        double4 r = red[tl];
        for (int k = 1; k < 4; ++k) {
            const double4 o = red[tl + k * kDirTargets];
            r.x += o.x; r.y += o.y; r.z += o.z; r.w += o.w;
        }
        part[(size_t)tgt * kDirChunks + blockIdx.y] = r;
    }
}

__global__ __launch_bounds__(256) void k_direct64_reduce(const double4 *part, int ns, double4 *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int k = 0; k < kDirChunks; ++k) {
        const double4 o = part[(size_t)t * kDirChunks + k];
        r.x += o.x; r.y += o.y; r.z += o.z; r.w += o.w;
    }
    out[t] = r;
}

// part must hold ns * 64 entries, out ns entries
hipError_t launch_direct64(const Dev &d, const int *idx, int ns, double4 *part, double4 *out, bool phi,
                           hipStream_t st)
{
    if (ns <= 0) return hipSuccess;
    const dim3 grid((ns + kDirTargets - 1) / kDirTargets, kDirChunks);
    if (phi)
        hipLaunchKernelGGL(k_direct64<true>, grid, dim3(256), 0, st, d, idx, ns, part);
    else
        hipLaunchKernelGGL(k_direct64<false>, grid, dim3(256), 0, st, d, idx, ns, part);
    hipLaunchKernelGGL(k_direct64_reduce, dim3((ns + 255) / 256), dim3(256), 0, st, (const double4 *)part, ns,
                       out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 sums over all bodies for the diagnostics: per block (grid-stride over
// bodies, fixed grid) the mass, momentum, kinetic energy, angular momentum
// about the origin and first moment of mass; the host adds the blocks in order.
// ---------------------------------------------------------------------------
constexpr int kMomTerms = 11; // m, p(3), ke, L(3), m x(3)

__global__ __launch_bounds__(256) void k_moments(Dev d, double *part)
{
    __shared__ double red[kMomTerms][256];
    double acc[kMomTerms];
#pragma unroll
    for (int k = 0; k < kMomTerms; ++k) acc[k] = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < d.n; i += kMomentBlocks * 256) {
        const float4 p = d.pm4[i];
        const float4 v = d.vel4[i];
        double x = p.x, y = p.y, z = p.z;
        if (d.fp64) {
            const double4 q = d.pos64[i];
            x = q.x; y = q.y; z = q.z;
        }
        const double m = p.w, vx = v.x, vy = v.y, vz = v.z;
        acc[0] += m;
        acc[1] += m * vx;
        acc[2] += m * vy;
        acc[3] += m * vz;
        acc[4] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        acc[5] += m * (y * vz - z * vy);
        acc[6] += m * (z * vx - x * vz);
        acc[7] += m * (x * vy - y * vx);
        acc[8] += m * x;
        acc[9] += m * y;
        acc[10] += m * z;
    }
#pragma unroll
    for (int k = 0; k < kMomTerms; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < kMomTerms; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < kMomTerms) part[(size_t)blockIdx.x * kMomTerms + threadIdx.x] = red[threadIdx.x][0];
}

hipError_t launch_moments(const Dev &d, double *part, hipStream_t st)
{
    hipLaunchKernelGGL(k_moments, dim3(kMomentBlocks), dim3(256), 0, st, d, part);
    return hipGetLastError();
}

} // namespace nb
