// nb_learned.hip -- the learned coarse-theta correction for gfx950.
//
// Coarse traversal. The wave-uniform traversal of the force stage (one wave
// = one vote group of 64 sorted bodies, a cell is taken when all 64 lanes
// accept it) with a coarser test. A lane accepts a cell of extent b (the
// largest distance of its bodies from its centre of mass, bounded in
// summarize) and mass m at s^2 = |d|^2 + eps^2 when
//     s^2 >= b^2 / theta_c^2 + eps^2           (extent test: |d| >= b / theta_c)
//     s^2 >= (m b^3)^(2/5) (alpha^2 |a_old|^2)^(-1/5)
//                                              (relative test: m b^3 / s^5, the
//                                               size of the first term left
//                                               out, within alpha |a_old|; its
//                                               two factors are rounded to FP32
//                                               once, per cell and per body)
// with a_old the body's acceleration of the previous step (the relative test
// is off for a lane while there is none). Every accepted cell contributes its
// monopole and its quadrupole under the softened kernel: with d = centre of
// mass - body and the cell's traceless second moment Q and trace T about its
// centre of mass,
//     a = m d / s^3 - Q d / s^5 + 5/2 (d.Q.d - T eps^2) d / s^7 .
// (The T eps^2 term is the part of the second-order expansion of the
// softened kernel that the traceless Q does not carry; it matters once cells
// sit within a few softening lengths.) Bodies interact as point masses. The
// training targets use the same kernel with the force stage's opening table
// (kOutAccel).
//
// Features. Per body, 24 numbers in a frame built from the body's own coarse
// acceleration a: e1 = a / |a|, e2 = the part of the quadrupole acceleration
// a_q perpendicular to e1, normalized, e3 = e1 x e2. Vectors are projected on
// the frame and divided by |a|, so the features do not change under rotations
// and under a common scaling of the masses. Per accepted cell the traversal
// only keeps the quadrupole part and the cell mass apart; the rest comes once
// per body from wave reductions.
//    0-2  acceleration from bodies (a_b)
//    3-4  a_q along e1, e2 (its e3 component is 0 by construction)
//    5-7  a minus the mean coarse acceleration of the group
//    8-10 offset of the body from its group's centroid, over the group radius
//   11-13 the group's mean a_b
//   14-16 the group's mean a_q
//   17-18 log(1 + accepted cells), log(1 + bodies met)
//   19-20 log(|a_q| / |a|), log(|a_b| / |a|)
//   21    accepted-cell mass / (accepted-cell mass + body mass met)
//   22    log(group radius / root radius)
//   23    cosine between the cell monopole part and a_b
//
// Network. 24 -> 64 -> 64 -> 3 with tanh-GELU. Inference runs in the epilogue
// of the coarse traversal itself: the wave's 64 bodies go through
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation) in the
// transposed form Z^T = W^T X^T, 16 bodies (one B-operand column block) at a
// time: 4 + 8 + 2 matrix instructions per block. The features reach the
// operand layout through LDS once; between layers, each lane's D registers
// already hold what it needs as the next B operand (the weights are stored
// with the matching order of k), and the three outputs return to their own
// lanes by shuffles. The corrected acceleration then goes straight into the
// body update, so the update overlaps with other waves' traversal. Training:
// FP32 kernels for the forward pass, the backward pass and Adam, with
// deterministic reductions.
#include "nb_device.h"

namespace nb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// numerics shared by inference and training
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t to_bf16(float f)
{
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// tanh(y) = 1 - 2 / (exp(2y) + 1) with the hardware reciprocal (1 ulp): the
// network is BF16 at inference, so an IEEE division buys nothing here;
// saturates correctly at both ends (rcp(inf) = 0, rcp(1) = 1)
__device__ __forceinline__ float tanh_fast(float y)
{
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * y) + 1.0f);
}

constexpr float kGeluC = 0.7978845608028654f; // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

__device__ __forceinline__ float gelu(float x)
{
    const float t = tanh_fast(kGeluC * (x + kGeluA * x * x * x));
    return 0.5f * x * (1.0f + t);
}

__device__ __forceinline__ float gelu_grad(float x)
{
    const float x2 = x * x;
    const float t = tanh_fast(kGeluC * (x + kGeluA * x2 * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * x2);
}

__device__ __forceinline__ float clipz(float z) { return fminf(fmaxf(z, -kClipZ), kClipZ); }

// Inference GELU in the logistic form of the same function,
// 0.5 x (1 + tanh(u)) = x / (1 + exp(-2 u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// with the hardware 2^x and reciprocal: 7 instructions. Limits: x -> +inf
// gives x, x -> -inf gives -0 (2^+inf = inf, 1 / inf = 0).
constexpr float kGeluK1 = -2.0f * kGeluC * 1.4426950408889634f;          // -2 c log2(e)
constexpr float kGeluK2 = -2.0f * kGeluC * kGeluA * 1.4426950408889634f; // -2 c a log2(e)

__device__ __forceinline__ float gelu_infer(float x)
{
    const float s = fmaf(kGeluK2, x * x, kGeluK1);
    return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(s * x));
}

// two floats to packed BF16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 h = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, h);
}

__device__ __forceinline__ bf16x8 load_bf16x8(const uint16_t *p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(p));
}

constexpr int kStackC = 768; // >= kStack; 768 words = 64 rows of 24 BF16 features
static_assert(kStackC >= kStack && kStackC * 4 >= 64 * kFeat * 2, "coarse stack doubles as the feature rows");

// ---------------------------------------------------------------------------
// BF16 matrix-core network (shared by the learned step and the test path)
//
// v_mfma_f32_16x16x32_bf16 operand layout (lane l, i16 = l % 16, g = l / 16):
//   A: row i16 of the 16-row block, k = 8 g .. 8 g + 7
//   B: column i16 of the 16-column block, k = 8 g .. 8 g + 7
//   D: column i16, rows 4 g .. 4 g + 3
// ---------------------------------------------------------------------------
// GELU of two accumulator blocks (their biases came in as the accumulators'
// initial values), packed to BF16
__device__ __forceinline__ bf16x8 pack_gelu(f32x4 lo, f32x4 hi)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 h = {pack_bf16(gelu_infer(lo[0]), gelu_infer(lo[1])), pack_bf16(gelu_infer(lo[2]), gelu_infer(lo[3])),
                     pack_bf16(gelu_infer(hi[0]), gelu_infer(hi[1])), pack_bf16(gelu_infer(hi[2]), gelu_infer(hi[3]))};
    return __builtin_bit_cast(bf16x8, h);
}

__device__ __forceinline__ f32x4 as_f32x4(float4 v) { return f32x4{v.x, v.y, v.z, v.w}; }

// Runs the wave's 64 rows of normalized BF16 features (fs: LDS, row r at
// fs + 24 r) through the network; lane r receives the residual of row r.
__device__ __forceinline__ void mlp_wave(const uint16_t *fs, const Mlp &m, int lane, float &y0, float &y1,
                                         float &y2)
{
    const int i16 = lane & 15, g = lane >> 4;
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    y0 = y1 = y2 = 0.f;
    // one 16-body column block at a time, not unrolled: the epilogue must not
    // raise the register count of the traversal it ends
#pragma unroll 1
    for (int cb = 0; cb < 4; ++cb) {
        const bf16x8 xb = g < 3 ? load_bf16x8(fs + (cb * 16 + i16) * kFeat + 8 * g)
                                : __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0});
        // each layer in two halves of 32 hidden rows (two 16-row blocks),
        // each half's activations packed before the next half starts: fewer
        // accumulators live at once
        const float4 *b1 = reinterpret_cast<const float4 *>(m.b1 + 4 * g);
        const float4 *b2 = reinterpret_cast<const float4 *>(m.b2 + 4 * g);
        bf16x8 h1[2];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            const f32x4 p0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_bf16x8(m.w1f + ((2 * hf) * 64 + lane) * 8), xb,
                                                                     as_f32x4(b1[8 * hf]), 0, 0, 0);
            const f32x4 p1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                load_bf16x8(m.w1f + ((2 * hf + 1) * 64 + lane) * 8), xb, as_f32x4(b1[8 * hf + 4]), 0, 0, 0);
            h1[hf] = pack_gelu(p0, p1);
            __builtin_amdgcn_sched_barrier(0);
        }
        bf16x8 h2[2];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            f32x4 p[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int rb = 2 * hf + q;
                p[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_bf16x8(m.w2f + ((rb * 2 + 0) * 64 + lane) * 8),
                                                               h1[0], as_f32x4(b2[4 * rb]), 0, 0, 0);
                p[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_bf16x8(m.w2f + ((rb * 2 + 1) * 64 + lane) * 8),
                                                               h1[1], p[q], 0, 0, 0);
            }
            h2[hf] = pack_gelu(p[0], p[1]);
            __builtin_amdgcn_sched_barrier(0);
        }
        const bf16x8 h2a = h2[0], h2b = h2[1];
        f32x4 d3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_bf16x8(m.w3f + lane * 8), h2a, z4, 0, 0, 0);
        d3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_bf16x8(m.w3f + (64 + lane) * 8), h2b, d3, 0, 0, 0);
        // outputs 0..2 of the column block's bodies sit in lanes 0..15 (g = 0)
        const float t0 = __shfl(d3[0], i16), t1 = __shfl(d3[1], i16), t2 = __shfl(d3[2], i16);
        if (g == cb) {
            y0 = t0 + m.b3[0];
            y1 = t1 + m.b3[1];
            y2 = t2 + m.b3[2];
        }
    }
}

// ---------------------------------------------------------------------------
// Coarse traversal
// ---------------------------------------------------------------------------
// lane k's value of x (k wave-uniform)
__device__ __forceinline__ float lane_read(float x, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}

#define NB_COARSE_PARAMS                                                                                     \
    Dev d, Learn L, const int *__restrict__ cchild, const float4 *__restrict__ cdata,                        \
        const float4 *__restrict__ cqa, const float4 *__restrict__ cqb, const float2 *__restrict__ cmac,     \
        const float *__restrict__ dqtab, const int *__restrict__ groups, int ngroups, float *__restrict__ xraw, \
        float4 *__restrict__ frame_out, float4 *__restrict__ acc_out, int fmode
#define NB_COARSE_PASS d, L, cchild, cdata, cqa, cqb, cmac, dqtab, groups, ngroups, xraw, frame_out, acc_out, fmode

// waves (vote groups) per block of the coarse traversal. The waves share
// nothing, and a block of one wave frees its slot as soon as its group is
// done, where a block of four holds a CU's share until its slowest group
// ends (learned force 3.48 -> 3.42 ms at 4M against four, 3.44 with two)
constexpr int kCoarseWaves = 1;

template <int OUT> __device__ __forceinline__ void coarse_traverse(NB_COARSE_PARAMS)
{
    constexpr bool kFeatOn = OUT == kOutLearned || OUT == kOutRaw;
    constexpr bool kGeo = OUT == kOutAccel; // the fine traversal: the force stage's opening table
    constexpr bool kLaneQ = true; // quadrupole records by vector loads, read back from LDS
    // a wave's traversal stack; after the traversal the learned step reuses it
    // for the wave's 64 rows of BF16 features (64 x 24 x 2 bytes)
    __shared__ __attribute__((aligned(16))) unsigned stk[kCoarseWaves][kStackC];
    // a popped cell's children's quadrupole records, (qa, qb) per child, for
    // the accepted cells' uniform LDS reads
    __shared__ float4 qst[kCoarseWaves][16];
    __shared__ float dq_s[kMaxDepth];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int n = d.n;
    Ctl *ctl = d.ctl;
    if (ctl->error) return;
    if (ctl->max_depth > kMaxDepth) {
        if (tid == 0) flag_error(ctl, kErrDepth);
        return;
    }
    if (kGeo)
        for (int i = tid; i < kMaxDepth; i += 64 * kCoarseWaves) dq_s[i] = dqtab[i];
    const int step = ctl->step;
    const float root_r = ctl->radius;
    __syncthreads();

    const int gslot = blockIdx.x * kCoarseWaves + wv;
    int g = gslot;
    if (groups) {
        if (gslot >= ngroups) return;
        g = groups[gslot];
    }
    const int pos = g * 64 + lane;
    const bool valid = pos >= 0 && pos < n;
    if (__ballot(valid) == 0ull) return; // wave-uniform
    // a lane past the last body sits at infinity: its d^2 is infinite, so it
    // accepts every cell and never holds up the vote (its sums are discarded)
    const float4 me = valid ? d.pm4[d.sorted[pos]] : make_float4(kFar, kFar, kFar, 0.f);
    const float eps2 = d.eps2;
    const int cap = d.cap;
    // the body's factor (alpha^2 |a_old|^2)^(-1/5) of the relative test's
    // threshold; 0 switches it off (no previous acceleration yet, as in the
    // force stage's kick, or alpha = 0)
    float a5 = 0.f;
    if (!kGeo && valid && step > 0 && d.alpha2_coarse > 0.f) {
        const float4 a0 = d.acc4[d.sorted[pos]];
        const float a2 = fmaf(a0.x, a0.x, fmaf(a0.y, a0.y, a0.z * a0.z));
        if (a2 > 0.f) a5 = (float)pow((double)(d.alpha2_coarse * a2), -0.2);
    }

    if (lane == 0) stk[wv][0] = 0; // root cell 0 at depth 0
    int sp = 1;
    // The fine traversal (training targets) sums each popped cell's terms into
    // a compensated total (s, k). The coarse one keeps three running sums --
    // bodies (b), cell monopoles (s; with the quadrupole when no features are
    // made) and cell quadrupoles (q) -- and adds them at the end: its error
    // bound is 1e-3, and the compensation cost 12 of its 27 instructions per
    // popped cell.
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    float bx = 0.f, by = 0.f, bz = 0.f;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    float smb = 0.f; // mass of the bodies met (the accepted cells hold the rest)
    int nc = 0, nbod = 0;
    long long ninter = 0, ncell = 0;
    int npop = 0;
    bool bad = false;
    while (sp > 0) {
        __builtin_amdgcn_wave_barrier();
        --sp;
        ++npop;
        const unsigned e = __builtin_amdgcn_readfirstlane(stk[wv][sp]);
        const int c = (int)(e >> 6);
        const int dep = (int)(e & 63u);
        const float thr_geo = kGeo ? dq_s[dep] : 0.f;
        const int *chp = cchild + (size_t)c * 8;
        const float4 *cdp = cdata + (size_t)c * 8;
        int ch[8];
        float4 q[8];
        float2 mc[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ch[k] = __builtin_amdgcn_readfirstlane(chp[k]);
            q[k] = cdp[k];
            if (!kGeo) mc[k] = cmac[(size_t)c * 8 + k];
        }
        // the children's quadrupole records, child k in lanes k, k + 8, ...:
        // two vector loads issued now, waited for at the first accepted cell
        // (which reads them from the wave's LDS), so their latency hides
        // behind the decisions and the bodies; each from the popped cell's
        // uniform base plus an unsigned lane offset (fewer 64-bit address
        // adds per pop)
        float4 qav, qbv;
        if constexpr (kLaneQ) {
            qav = (cqa + (size_t)c * 8)[(unsigned)lane & 7u];
            qbv = (cqb + (size_t)c * 8)[(unsigned)lane & 7u];
        }
        // First the opening decisions, wave-uniform bit masks of the children
        // (children are compacted to the front: past the first negative entry
        // every one is negative). A cell: the opening test of this traversal.
        // The coarse one takes the extent test first (threshold b^2 /
        // theta_c^2 + eps^2 from summarize) and the relative test (threshold
        // the product of the cell's and the body's factors) only when every
        // lane passed it: both votes are wave-uniform, so the short circuit is
        // a scalar branch.
        // Lane masks come straight from the compares (3 = ordered >=, 5 =
        // ordered <=). An opened cell's stack entry goes into lane `npush` of
        // `push` (v_writelane: no compare, no branch on the lane); all of them
        // are stored after the loop.
        unsigned bodies = 0, cells = 0;
        int npush = 0;
        bool badv = false;
        unsigned push = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = ch[k];
            if (v < 0) continue;
            if (v < n) {
                bodies |= 1u << k;
                continue;
            }
            const float dx = q[k].x - me.x, dy = q[k].y - me.y, dz = q[k].z - me.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            const uint64_t exec = __builtin_amdgcn_read_exec();
            bool take = __builtin_amdgcn_fcmpf(d2, kGeo ? thr_geo : mc[k].x, 3) == exec;
            if (!kGeo && take) take = __builtin_amdgcn_fcmpf(mc[k].y * a5, d2, 5) == exec;
            if (take) {
                cells |= 1u << k;
            } else {
                push = lane_write(push, ((unsigned)(v - n) << 6) | (unsigned)(dep + 1), npush);
                ++npush;
                badv |= v - n >= cap;
            }
        }
        // Then the interactions, each kind in child order, each branch only
        // around its own work, so the running sums are updated in place (a
        // branch that both accumulates and pushes makes the compiler copy
        // them at every child): bodies, cell monopoles and cell quadrupoles
        // summed separately (the features need the parts). The learned step
        // adds its terms to its three running sums directly; the fine
        // traversal sums a popped cell's terms from zero for its compensated
        // add, and the coarse one without features (one sum of monopole and
        // quadrupole) from zero for its running sums, which keeps it within
        // 1e-5 of the oracle's FP64 sums. d is worked out again, with the
        // same operations as for the test.
        float pbx, pby, pbz, pmx, pmy, pmz, pqx, pqy, pqz;
        if constexpr (kGeo || !kFeatOn) {
            pbx = pby = pbz = pmx = pmy = pmz = pqx = pqy = pqz = 0.f;
        } else {
            pbx = bx;
            pby = by;
            pbz = bz;
            pmx = sx;
            pmy = sy;
            pmz = sz;
            pqx = qx;
            pqy = qy;
            pqz = qz;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!((bodies >> k) & 1u)) continue;
            const float4 qk = q[k];
            const float dx = qk.x - me.x, dy = qk.y - me.y, dz = qk.z - me.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            const float ri = __builtin_amdgcn_rsqf(d2);
            const float f = qk.w * (ri * ri * ri);
            pbx = fmaf(dx, f, pbx);
            pby = fmaf(dy, f, pby);
            pbz = fmaf(dz, f, pbz);
            if (kFeatOn) smb += qk.w;
        }
        // the children's quadrupole records into this wave's LDS rows (the
        // vector loads have had the decisions and the bodies to arrive): an
        // accepted cell reads its two with uniform LDS reads, no VALU
        if constexpr (kLaneQ) {
            if (cells) {
                if (lane < 8) {
                    qst[wv][2 * lane] = qav;
                    qst[wv][2 * lane + 1] = qbv;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (!((cells >> k) & 1u)) continue;
            const float4 qk = q[k];
            const float dx = qk.x - me.x, dy = qk.y - me.y, dz = qk.z - me.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            float4 qa, qb; // (Qxx, Qyy, Qzz, T eps^2), (Qxy, Qxz, Qyz, 0)
            if constexpr (kLaneQ) {
                qa = qst[wv][2 * k];
                qb = qst[wv][2 * k + 1];
            } else {
                qa = cqa[(size_t)c * 8 + k];
                qb = cqb[(size_t)c * 8 + k];
            }
            const float ri = __builtin_amdgcn_rsqf(d2);
            const float r2i = ri * ri;
            const float r3 = r2i * ri;
            const float f = qk.w * r3;
            const float Qx = fmaf(qa.x, dx, fmaf(qb.x, dy, qb.y * dz));
            const float Qy = fmaf(qb.x, dx, fmaf(qa.y, dy, qb.z * dz));
            const float Qz = fmaf(qb.y, dx, fmaf(qb.z, dy, qa.z * dz));
            const float dQd = fmaf(Qx, dx, fmaf(Qy, dy, Qz * dz));
            const float r5 = r3 * r2i;
            if (kFeatOn) {
                // the quadrupole part as r5 (u d - Q d), u = 5/2 (d.Q.d - T eps^2) / s^2
                const float u = (dQd - qa.w) * (2.5f * r2i);
                pmx = fmaf(dx, f, pmx);
                pmy = fmaf(dy, f, pmy);
                pmz = fmaf(dz, f, pmz);
                pqx = fmaf(r5, fmaf(u, dx, -Qx), pqx);
                pqy = fmaf(r5, fmaf(u, dy, -Qy), pqy);
                pqz = fmaf(r5, fmaf(u, dz, -Qz), pqz);
            } else {
                const float c2 = (dQd - qa.w) * ((2.5f * r5) * r2i);
                // monopole and quadrupole radial terms share one factor
                const float h = f + c2;
                pmx = fmaf(dx, h, fmaf(-r5, Qx, pmx));
                pmy = fmaf(dy, h, fmaf(-r5, Qy, pmy));
                pmz = fmaf(dz, h, fmaf(-r5, Qz, pmz));
            }
        }
        const int nb_here = __popc(bodies), nc_here = __popc(cells);
        const int cnt = nb_here + nc_here;
        ncell += nc_here;
        if (kFeatOn) {
            nbod += nb_here;
            nc += nc_here;
        }
        // the opened children in one store, in child order (the last one is
        // popped first); the tree is at most kMaxDepth deep and the pool holds
        // `cap` cells: anything else is corrupt
        if (npush) {
            if (sp + npush > kStackC || badv || dep + 1 >= kMaxDepth) {
                bad = true;
            } else {
                if (lane < npush) stk[wv][sp + lane] = push;
                sp += npush;
            }
        }
        if constexpr (kGeo) {
            // the training targets: compensated
            kahan_add(sx, kx, (pbx + pmx) + pqx);
            kahan_add(sy, ky, (pby + pmy) + pqy);
            kahan_add(sz, kz, (pbz + pmz) + pqz);
        } else if constexpr (!kFeatOn) {
            sx += pmx;
            sy += pmy;
            sz += pmz;
            bx += pbx;
            by += pby;
            bz += pbz;
        } else {
            // the coarse force: its parts summed apart (the features need them)
            sx = pmx;
            sy = pmy;
            sz = pmz;
            bx = pbx;
            by = pby;
            bz = pbz;
            qx = pqx;
            qy = pqy;
            qz = pqz;
        }
        ninter += cnt;
        if (bad) break;
    }
    if (bad && lane == 0) flag_error(ctl, kErrStack);
    if constexpr (OUT == kOutLearned || OUT == kOutFinish) {
        // the group's cost for the next launch order, in interactions: a
        // popped cell (its loads and up to eight tests) weighs about three
        const long long cost = ninter + 3ll * npop;
        if (lane == 0) d.gcost[g] = cost < 0x7fffffffll ? (int)cost : 0x7fffffff;
    }
    const float ax = kGeo ? sx - kx : (bx + sx) + qx;
    const float ay = kGeo ? sy - ky : (by + sy) + qy;
    const float az = kGeo ? sz - kz : (bz + sz) + qz;
    const int slot = gslot * 64 + lane;
    if constexpr (OUT == kOutAccel) {
        if (valid) acc_out[slot] = make_float4(ax, ay, az, 0.f);
    } else if constexpr (OUT == kOutFinish) {
        if (valid) finish_body(d, pos, me, ax, ay, az, fmode, step);
    } else {
        // group centroid and radius, mean acceleration and mean body part
        // (wave reductions over the valid lanes)
        const float nv = (float)__popcll(__ballot(valid));
        float gx = valid ? me.x : 0.f, gy = valid ? me.y : 0.f, gz = valid ? me.z : 0.f;
        float gax = valid ? ax : 0.f, gay = valid ? ay : 0.f, gaz = valid ? az : 0.f;
        float gbx = valid ? bx : 0.f, gby = valid ? by : 0.f, gbz = valid ? bz : 0.f;
        float gqx = valid ? qx : 0.f, gqy = valid ? qy : 0.f, gqz = valid ? qz : 0.f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            gqx += __shfl_xor(gqx, o);
            gqy += __shfl_xor(gqy, o);
            gqz += __shfl_xor(gqz, o);
            gx += __shfl_xor(gx, o);
            gy += __shfl_xor(gy, o);
            gz += __shfl_xor(gz, o);
            gax += __shfl_xor(gax, o);
            gay += __shfl_xor(gay, o);
            gaz += __shfl_xor(gaz, o);
            gbx += __shfl_xor(gbx, o);
            gby += __shfl_xor(gby, o);
            gbz += __shfl_xor(gbz, o);
        }
        // the features end as BF16 network inputs: hardware reciprocal,
        // square root and logarithm (about 1 ulp) instead of the IEEE sequences
        const float inv_nv = __builtin_amdgcn_rcpf(nv);
        gx *= inv_nv;
        gy *= inv_nv;
        gz *= inv_nv;
        gax *= inv_nv;
        gay *= inv_nv;
        gaz *= inv_nv;
        gbx *= inv_nv;
        gby *= inv_nv;
        gbz *= inv_nv;
        gqx *= inv_nv;
        gqy *= inv_nv;
        gqz *= inv_nv;
        const float ox = me.x - gx, oy = me.y - gy, oz = me.z - gz;
        float rg = valid ? __builtin_amdgcn_sqrtf(ox * ox + oy * oy + oz * oz) : 0.f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) rg = fmaxf(rg, __shfl_xor(rg, o));

        float f[kFeat];
#pragma unroll
        for (int i = 0; i < kFeat; ++i) f[i] = 0.f;
        const float A = __builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az);
        float e2x = 0.f, e2y = 0.f, e2z = 0.f;
        if (A > 0.f) {
            const float inv = __builtin_amdgcn_rcpf(A);
            const float e1x = ax * inv, e1y = ay * inv, e1z = az * inv;
            const float qe1 = qx * e1x + qy * e1y + qz * e1z;
            float wx = qx - qe1 * e1x, wy = qy - qe1 * e1y, wz = qz - qe1 * e1z;
            float W = __builtin_amdgcn_sqrtf(wx * wx + wy * wy + wz * wz);
            if (!(W > 0.f)) {
                // no quadrupole direction: the coordinate axis least aligned with e1
                const float fx = fabsf(e1x), fy = fabsf(e1y), fz = fabsf(e1z);
                const float tx = (fx <= fy && fx <= fz) ? 1.f : 0.f;
                const float ty = (tx == 0.f && fy <= fz) ? 1.f : 0.f;
                const float tz = (tx == 0.f && ty == 0.f) ? 1.f : 0.f;
                const float te1 = tx * e1x + ty * e1y + tz * e1z;
                wx = tx - te1 * e1x;
                wy = ty - te1 * e1y;
                wz = tz - te1 * e1z;
                W = __builtin_amdgcn_sqrtf(wx * wx + wy * wy + wz * wz);
            }
            const float iW = __builtin_amdgcn_rcpf(W);
            e2x = wx * iW;
            e2y = wy * iW;
            e2z = wz * iW;
            const float e3x = e1y * e2z - e1z * e2y;
            const float e3y = e1z * e2x - e1x * e2z;
            const float e3z = e1x * e2y - e1y * e2x;
#define NB_LOC(X, Y, Z, I)                                                                         \
    f[I] = ((X) * e1x + (Y) * e1y + (Z) * e1z) * inv;                                             \
    f[(I) + 1] = ((X) * e2x + (Y) * e2y + (Z) * e2z) * inv;                                       \
    f[(I) + 2] = ((X) * e3x + (Y) * e3y + (Z) * e3z) * inv
            NB_LOC(bx, by, bz, 0);
            f[3] = qe1 * inv;
            f[4] = (qx * e2x + qy * e2y + qz * e2z) * inv;
            NB_LOC(ax - gax, ay - gay, az - gaz, 5);
            if (rg > 0.f) {
                const float ir = __builtin_amdgcn_rcpf(rg);
                f[8] = (ox * e1x + oy * e1y + oz * e1z) * ir;
                f[9] = (ox * e2x + oy * e2y + oz * e2z) * ir;
                f[10] = (ox * e3x + oy * e3y + oz * e3z) * ir;
            }
            NB_LOC(gbx, gby, gbz, 11);
            NB_LOC(gqx, gqy, gqz, 14);
            f[17] = __logf(1.0f + (float)nc); // counts below 2^24: 1 + n is exact
            f[18] = __logf(1.0f + (float)nbod);
            const float qn = __builtin_amdgcn_sqrtf(qx * qx + qy * qy + qz * qz);
            const float bn = __builtin_amdgcn_sqrtf(bx * bx + by * by + bz * bz);
            f[19] = __logf(fmaxf(qn * inv, 1e-30f));
            f[20] = __logf(fmaxf(bn * inv, 1e-30f));
            // accepted-cell mass over the mass met: every body is met once, as a
            // body or inside an accepted cell, so the cells hold the total
            // mass (the root's) minus the bodies' (one sum fewer in the loop)
            const float mtot = (float)d.com64[0].w;
            f[21] = mtot > 0.f ? (mtot - smb) * __builtin_amdgcn_rcpf(mtot) : 0.f;
            f[22] = __logf(fmaxf(rg * __builtin_amdgcn_rcpf(root_r), 1e-30f));
            const float mx = ax - bx - qx, my = ay - by - qy, mz = az - bz - qz;
            const float mn = __builtin_amdgcn_sqrtf(mx * mx + my * my + mz * mz);
            f[23] = (mn > 0.f && bn > 0.f) ? (mx * bx + my * by + mz * bz) * __builtin_amdgcn_rcpf(mn * bn) : 0.f;
#undef NB_LOC
        }
        if constexpr (OUT == kOutRaw) {
            if (valid) {
                float *o = xraw + (size_t)slot * kFeat;
#pragma unroll
                for (int i = 0; i < kFeat; i += 4)
                    *reinterpret_cast<float4 *>(o + i) = make_float4(f[i], f[i + 1], f[i + 2], f[i + 3]);
                frame_out[2 * (size_t)slot] = make_float4(ax, ay, az, A);
                frame_out[2 * (size_t)slot + 1] = make_float4(A * e2x, A * e2y, A * e2z, 0.f);
            }
        } else {
            // the learned step: normalized BF16 features into this wave's LDS
            // rows, the network, the correction, the body update
            uint16_t *fs = reinterpret_cast<uint16_t *>(stk[wv]);
            uint32_t w[kFeat / 2];
#pragma unroll
            for (int i = 0; i < kFeat; i += 2) {
                const uint32_t h = pack_bf16(clipz((f[i] - L.mlp.xmu[i]) * L.mlp.xrs[i]),
                                             clipz((f[i + 1] - L.mlp.xmu[i + 1]) * L.mlp.xrs[i + 1]));
                w[i / 2] = valid ? h : 0u;
            }
            uint4 *o = reinterpret_cast<uint4 *>(fs + lane * kFeat);
            o[0] = make_uint4(w[0], w[1], w[2], w[3]);
            o[1] = make_uint4(w[4], w[5], w[6], w[7]);
            o[2] = make_uint4(w[8], w[9], w[10], w[11]);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            float y0, y1, y2;
            mlp_wave(fs, L.mlp, lane, y0, y1, y2);
            // residual = |a| (y0 e1 + y1 e2 + y2 e3)
            float rx = ax, ry = ay, rz = az;
            if (A > 0.f) {
                const float iA = __builtin_amdgcn_rcpf(A);
                const float e1x = ax * iA, e1y = ay * iA, e1z = az * iA;
                const float e3x = e1y * e2z - e1z * e2y, e3y = e1z * e2x - e1x * e2z, e3z = e1x * e2y - e1y * e2x;
                rx += A * (y0 * e1x + y1 * e2x + y2 * e3x);
                ry += A * (y0 * e1y + y1 * e2y + y2 * e3y);
                rz += A * (y0 * e1z + y1 * e2z + y2 * e3z);
            }
            if (valid) finish_body(d, pos, me, rx, ry, rz, fmode, step);
        }
    }
    // device interaction counter: one atomic per wave
    const unsigned long long nvalid = (unsigned long long)__popcll(__ballot(valid));
    if (lane == 0 && ninter && nvalid) {
        __hip_atomic_fetch_add(&ctl->interactions, (unsigned long long)ninter * nvalid, __ATOMIC_RELAXED,
                               AGENT);
        __hip_atomic_fetch_add(&ctl->cell_interactions, (unsigned long long)ncell * nvalid, __ATOMIC_RELAXED,
                               AGENT);
    }
}

// The kernels. Six waves per SIMD at least for the network's epilogue (80
// registers); the coarse force without it asks for seven, which its 68
// registers allow and which changes its scalar register budget and code
// (0.12-0.17 ms less at 4M; the learned kernel at seven spills in its
// epilogue and is slower).
template <int OUT>
__global__ __launch_bounds__(64 * kCoarseWaves) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_coarse64(NB_COARSE_PARAMS)
{
    coarse_traverse<OUT>(NB_COARSE_PASS);
}

template <>
__global__ __launch_bounds__(64 * kCoarseWaves) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_coarse64<kOutFinish>(NB_COARSE_PARAMS)
{
    coarse_traverse<kOutFinish>(NB_COARSE_PASS);
}

// One block: a histogram of the groups' cost classes (1 + floor(log2 cost),
// 0 for no cost yet), then the groups by class, the most expensive first.
// Inside a class the groups keep roughly their sorted (spatial) order, so
// neighbouring waves still share tree cells in the caches; measured at 4M
// clustered bodies, classes of one octave beat four classes per octave
// (139.9 against 139.4 steps/s; 137.3 in sorted order). The order inside a
// class is not fixed, and the forces do not depend on the order in which
// groups run.
__device__ __forceinline__ int cost_key(int c)
{
    return c <= 0 ? 0 : 32 - __clz(c); // 1 + floor(log2 c) <= 31 < kCostKeys
}

// Claims h[key] + (rank of this lane among the wave's active lanes with the
// same key) with one LDS atomic per distinct key in the wave (neighbouring
// groups mostly share a class, so that is one or two atomics, not 64). Every
// lane of the wave calls it; the loop runs once per distinct key.
__device__ __forceinline__ int wave_claim(int *h, int key, bool active)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long rem = __ballot(active);
    int slot = 0;
    while (rem) {
        const int leader = __ffsll((long long)rem) - 1;
        const int k = __shfl(key, leader);
        const bool mine = active && key == k;
        const unsigned long long m = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&h[k], __popcll(m));
        base = __shfl(base, leader);
        if (mine) slot = base + __popcll(m & below);
        rem &= ~m;
    }
    return slot;
}

// Each pass reads the costs kOrderBatch per thread at a time, all loads in
// flight together (one block cannot hide a load's latency otherwise: the
// first version, one load per round trip, took 103 us at 4M bodies).
constexpr int kOrderBatch = 16;

__global__ __launch_bounds__(1024) void k_group_order(const int *__restrict__ cost, int ng, int *__restrict__ order)
{
    __shared__ int h[kCostKeys];
    const int tid = threadIdx.x;
    if (tid < kCostKeys) h[tid] = 0;
    __syncthreads();
    for (int base = 0; base < ng; base += 1024 * kOrderBatch) { // uniform bound: whole waves call wave_claim
        int key[kOrderBatch];
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j) {
            const int g = base + j * 1024 + tid;
            key[j] = g < ng ? cost_key(cost[g]) : -1;
        }
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j) (void)wave_claim(h, key[j], key[j] >= 0);
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int k = kCostKeys - 1; k >= 0; --k) {
            const int t = h[k];
            h[k] = acc;
            acc += t;
        }
    }
    __syncthreads();
    // the same costs as the histogram read (the engine runs this kernel after
    // the pass that wrote them and holds the next pass until it is done), so
    // every group lands in [0, ng) exactly once
    for (int base = 0; base < ng; base += 1024 * kOrderBatch) {
        int key[kOrderBatch];
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j) {
            const int g = base + j * 1024 + tid;
            key[j] = g < ng ? cost_key(cost[g]) : -1;
        }
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j) {
            const int slot = wave_claim(h, key[j], key[j] >= 0);
            if (key[j] >= 0) order[slot] = base + j * 1024 + tid;
        }
    }
}

hipError_t launch_group_order(const Dev &d, hipStream_t st)
{
    const int ng = (d.n + 63) / 64;
    if (ng <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_group_order, dim3(1), dim3(1024), 0, st, (const int *)d.gcost, ng, d.gorder);
    return hipGetLastError();
}

hipError_t launch_coarse(const Dev &d, const Learn &L, int out, const float *dqtab, const int *groups,
                         int ngroups, float *xraw, float4 *frame_out, float4 *acc_out, int fmode,
                         hipStream_t st)
{
    const int waves = groups ? ngroups : (d.n + 63) / 64;
    if (waves <= 0) return hipSuccess;
    const dim3 grid((waves + kCoarseWaves - 1) / kCoarseWaves), block(64 * kCoarseWaves);
#define NB_COARSE_ARGS                                                                             \
    d, L, (const int *)d.cchild, (const float4 *)d.cdata, (const float4 *)d.cqa, (const float4 *)d.cqb, \
        (const float2 *)d.cmac, dqtab, groups, ngroups, xraw, frame_out, acc_out, fmode
    switch (out) {
    case kOutLearned: hipLaunchKernelGGL(k_coarse64<kOutLearned>, grid, block, 0, st, NB_COARSE_ARGS); break;
    case kOutRaw: hipLaunchKernelGGL(k_coarse64<kOutRaw>, grid, block, 0, st, NB_COARSE_ARGS); break;
    case kOutAccel: hipLaunchKernelGGL(k_coarse64<kOutAccel>, grid, block, 0, st, NB_COARSE_ARGS); break;
    case kOutFinish: hipLaunchKernelGGL(k_coarse64<kOutFinish>, grid, block, 0, st, NB_COARSE_ARGS); break;
    default: return hipErrorInvalidValue;
    }
#undef NB_COARSE_ARGS
    return hipGetLastError();
}

// Test path: raw features -> normalized BF16 rows in LDS -> network -> y.
__global__ __launch_bounds__(256) void k_mlp_rows(Mlp m, const float *__restrict__ x, int rows,
                                                  float *__restrict__ y)
{
    __shared__ __attribute__((aligned(16))) uint16_t fs[4][64 * kFeat];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row = (blockIdx.x * 4 + wv) * 64 + lane;
#pragma unroll
    for (int i = 0; i < kFeat; ++i)
        fs[wv][lane * kFeat + i] = row < rows ? to_bf16(clipz((x[(size_t)row * kFeat + i] - m.xmu[i]) * m.xrs[i])) : 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    float y0, y1, y2;
    mlp_wave(fs[wv], m, lane, y0, y1, y2);
    if (row < rows) {
        y[(size_t)row * 3 + 0] = y0;
        y[(size_t)row * 3 + 1] = y1;
        y[(size_t)row * 3 + 2] = y2;
    }
}

hipError_t launch_mlp_rows(const Learn &L, const float *xraw, int rows, float *y, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_rows, dim3((rows + 255) / 256), dim3(256), 0, st, L.mlp, xraw, rows, y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Training pairs
// ---------------------------------------------------------------------------
__global__ void k_make_target(const float4 *__restrict__ frame, const float4 *__restrict__ afine, int rows,
                              float *__restrict__ t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const float4 f0 = frame[2 * (size_t)i], f1 = frame[2 * (size_t)i + 1];
    const float4 af = afine[i];
    float t0 = 0.f, t1 = 0.f, t2 = 0.f;
    if (f0.w > 0.f) {
        const float inv = 1.0f / f0.w;
        const float e1x = f0.x * inv, e1y = f0.y * inv, e1z = f0.z * inv;
        const float e2x = f1.x * inv, e2y = f1.y * inv, e2z = f1.z * inv;
        const float e3x = e1y * e2z - e1z * e2y, e3y = e1z * e2x - e1x * e2z, e3z = e1x * e2y - e1y * e2x;
        const float rx = af.x - f0.x, ry = af.y - f0.y, rz = af.z - f0.z;
        t0 = (rx * e1x + ry * e1y + rz * e1z) * inv;
        t1 = (rx * e2x + ry * e2y + rz * e2z) * inv;
        t2 = (rx * e3x + ry * e3y + rz * e3z) * inv;
    }
    t[3 * (size_t)i + 0] = t0;
    t[3 * (size_t)i + 1] = t1;
    t[3 * (size_t)i + 2] = t2;
}

hipError_t launch_make_target(const float4 *frame, const float4 *afine, int rows, float *t, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_make_target, dim3((rows + 255) / 256), dim3(256), 0, st, frame, afine, rows, t);
    return hipGetLastError();
}

// one block per column; FP64 sums in a fixed order
__global__ __launch_bounds__(256) void k_colstats(const float *__restrict__ x, int rows, int cols,
                                                  float *__restrict__ mu, float *__restrict__ sd)
{
    __shared__ double s1[256], s2[256];
    const int c = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (int r = threadIdx.x; r < rows; r += 256) {
        const double v = (double)x[(size_t)r * cols + c];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = s1[0] / rows;
        const double var = fmax(s2[0] / rows - m * m, 0.0);
        const double s = sqrt(var);
        mu[c] = (float)m;
        sd[c] = s > 1e-12 ? (float)s : 1.0f;
    }
}

__global__ void k_apply_standardize(float *__restrict__ x, int rows, int cols, const float *__restrict__ mu,
                                    const float *__restrict__ sd)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const int c = (int)(i % cols);
    x[i] = clipz((x[i] - mu[c]) / sd[c]);
}

hipError_t launch_standardize(float *x, int rows, int cols, float *mu, float *sd, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_colstats, dim3(cols), dim3(256), 0, st, x, rows, cols, mu, sd);
    const size_t total = (size_t)rows * cols;
    hipLaunchKernelGGL(k_apply_standardize, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, x, rows,
                       cols, (const float *)mu, (const float *)sd);
    return hipGetLastError();
}

__global__ void k_gather_acc(const float4 *__restrict__ acc, const int *__restrict__ idx, int ns,
                             float4 *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) out[i] = acc[idx[i]];
}

hipError_t launch_gather_acc(const Dev &d, const int *idx, int ns, float4 *out, hipStream_t st)
{
    if (ns <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_acc, dim3((ns + 255) / 256), dim3(256), 0, st, (const float4 *)d.acc4, idx, ns,
                       out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP32 training: forward, backward, Adam
// ---------------------------------------------------------------------------
__device__ __forceinline__ int batch_row(uint64_t seed, int it, int r, int S)
{
    if (seed == 0) return r; // the whole dataset in order
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (((uint64_t)it << 32) + (uint64_t)r + 1ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (int)(z % (uint64_t)S);
}

constexpr int kRowsPerBlock = 16;

// Z = X W + b, H = gelu(Z) for 16 rows per block; K inputs (24 or 64), 64 outputs.
// gather: X rows come from the dataset through batch_row.
template <int K, bool GATHER>
__global__ __launch_bounds__(256) void k_fwd_hidden(const float *__restrict__ X, uint64_t seed, int it, int S,
                                                    const float *__restrict__ W, const float *__restrict__ b,
                                                    int B, float *__restrict__ Z, float *__restrict__ H)
{
    __shared__ float w[K * 64];
    __shared__ float xin[kRowsPerBlock][K];
    const int t = threadIdx.x;
    for (int i = t; i < K * 64; i += 256) w[i] = W[i];
    const int r0 = blockIdx.x * kRowsPerBlock;
    for (int i = t; i < kRowsPerBlock * K; i += 256) {
        const int rl = i / K, k = i % K;
        const int r = r0 + rl;
        float v = 0.f;
        if (r < B) v = GATHER ? X[(size_t)batch_row(seed, it, r, S) * K + k] : X[(size_t)r * K + k];
        xin[rl][k] = v;
    }
    __syncthreads();
    const int j = t & 63;
    const float bj = b[j];
    for (int rl = t >> 6; rl < kRowsPerBlock; rl += 4) {
        const int r = r0 + rl;
        if (r >= B) break;
        float z = bj;
#pragma unroll 8
        for (int k = 0; k < K; ++k) z = fmaf(xin[rl][k], w[k * 64 + j], z);
        Z[(size_t)r * 64 + j] = z;
        H[(size_t)r * 64 + j] = gelu(z);
    }
}

// Y = H2 W3 + b3; dY = 2 (Y - T) / (3 B); per-block sum of (Y - T)^2
__global__ __launch_bounds__(256) void k_fwd_out(const float *__restrict__ H2, const float *__restrict__ W3,
                                                 const float *__restrict__ b3, const float *__restrict__ T,
                                                 uint64_t seed, int it, int S, int B, float *__restrict__ dY,
                                                 float *__restrict__ loss_part)
{
    __shared__ float w[64 * 3];
    __shared__ float red[256];
    const int t = threadIdx.x;
    if (t < 64 * 3) w[t] = W3[t];
    __syncthreads();
    const int r = blockIdx.x * 256 + t;
    float l = 0.f;
    if (r < B) {
        float y0 = b3[0], y1 = b3[1], y2 = b3[2];
        const float *h = H2 + (size_t)r * 64;
        for (int k = 0; k < 64; ++k) {
            const float hk = h[k];
            y0 = fmaf(hk, w[k * 3 + 0], y0);
            y1 = fmaf(hk, w[k * 3 + 1], y1);
            y2 = fmaf(hk, w[k * 3 + 2], y2);
        }
        const float *tt = T + (size_t)batch_row(seed, it, r, S) * 3;
        const float e0 = y0 - tt[0], e1 = y1 - tt[1], e2 = y2 - tt[2];
        const float sc = 2.0f / (3.0f * (float)B);
        dY[(size_t)r * 3 + 0] = sc * e0;
        dY[(size_t)r * 3 + 1] = sc * e1;
        dY[(size_t)r * 3 + 2] = sc * e2;
        l = e0 * e0 + e1 * e1 + e2 * e2;
    }
    red[t] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) loss_part[blockIdx.x] = red[0];
}

// dZ = (dNext W^T) * gelu'(Z): W is [64][N] (N = 3 or 64); 16 rows per block
template <int N>
__global__ __launch_bounds__(256) void k_back_hidden(const float *__restrict__ dNext, const float *__restrict__ W,
                                                     const float *__restrict__ Z, int B, float *__restrict__ dZ)
{
    __shared__ float wt[N][65]; // wt[j][k] = W[k][j], padded
    __shared__ float dn[kRowsPerBlock][N];
    const int t = threadIdx.x;
    for (int i = t; i < 64 * N; i += 256) {
        const int k = i / N, j = i % N;
        wt[j][k] = W[i];
    }
    const int r0 = blockIdx.x * kRowsPerBlock;
    for (int i = t; i < kRowsPerBlock * N; i += 256) {
        const int rl = i / N, j = i % N;
        const int r = r0 + rl;
        dn[rl][j] = r < B ? dNext[(size_t)r * N + j] : 0.f;
    }
    __syncthreads();
    const int k = t & 63;
    for (int rl = t >> 6; rl < kRowsPerBlock; rl += 4) {
        const int r = r0 + rl;
        if (r >= B) break;
        float s = 0.f;
#pragma unroll 8
        for (int j = 0; j < N; ++j) s = fmaf(dn[rl][j], wt[j][k], s);
        dZ[(size_t)r * 64 + k] = s * gelu_grad(Z[(size_t)r * 64 + k]);
    }
}

// Partial weight gradients: block (k, ch) sums rows [ch B / C, (ch + 1) B / C)
// of A[r][k] dZ[r][j] (k < K), or of dZ[r][j] for the bias (k = K); each
// thread takes every 4th row and the 4 partial sums are added in a fixed
// order. k_wgrad_reduce then adds the C chunks in order: deterministic.
template <bool GATHER>
__global__ __launch_bounds__(256) void k_wgrad(const float *__restrict__ A, int lda, uint64_t seed, int it, int S,
                                               const float *__restrict__ dZ, int N, int B, int K,
                                               float *__restrict__ part)
{
    __shared__ float red[4][64];
    const int t = threadIdx.x;
    const int j = t & 63, rq = t >> 6;
    const int k = blockIdx.x, ch = blockIdx.y;
    const int r0 = (int)((long long)B * ch / kGradChunks), r1 = (int)((long long)B * (ch + 1) / kGradChunks);
    float acc = 0.f;
    if (j < N) {
        for (int r = r0 + rq; r < r1; r += 4) {
            float a = 1.0f;
            if (k < K) a = GATHER ? A[(size_t)batch_row(seed, it, r, S) * lda + k] : A[(size_t)r * lda + k];
            acc = fmaf(a, dZ[(size_t)r * N + j], acc);
        }
    }
    red[rq][j] = acc;
    __syncthreads();
    if (rq == 0 && j < N)
        part[((size_t)ch * (K + 1) + k) * N + j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
}

__global__ void k_wgrad_reduce(const float *__restrict__ part, int K, int N, float *__restrict__ gW,
                               float *__restrict__ gb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (K + 1) * N) return;
    float v = 0.f;
    for (int ch = 0; ch < kGradChunks; ++ch) v += part[(size_t)ch * (K + 1) * N + i];
    const int k = i / N, j = i % N;
    if (k < K)
        gW[k * N + j] = v;
    else
        gb[j] = v;
}

__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                       float *__restrict__ V, int np, float lr, float b1, float b2, float eps, float c1, float c2,
                       const float *__restrict__ loss_part, int nparts, int B, float *__restrict__ loss_out,
                       bool step)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        float s = 0.f;
        for (int p = 0; p < nparts; ++p) s += loss_part[p];
        *loss_out = s / (3.0f * (float)B);
    }
    if (!step || i >= np) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

hipError_t launch_train_iter(const Train &t, int B, int it, uint64_t seed, bool adam, float lr, float beta1,
                             float beta2, float eps, float c1, float c2, hipStream_t st)
{
    if (B <= 0) return hipErrorInvalidValue;
    const dim3 blk(256);
    const dim3 gh((B + kRowsPerBlock - 1) / kRowsPerBlock);
    const int nparts = (B + 255) / 256;
    const float *P = t.P;
    float *G = t.G;
    hipLaunchKernelGGL((k_fwd_hidden<kFeat, true>), gh, blk, 0, st, t.X, seed, it, t.S, P, P + kOffB1, B, t.Z1, t.H1);
    hipLaunchKernelGGL((k_fwd_hidden<kHid, false>), gh, blk, 0, st, (const float *)t.H1, seed, it, t.S,
                       P + kOffW2, P + kOffB2, B, t.Z2, t.H2);
    hipLaunchKernelGGL(k_fwd_out, dim3(nparts), blk, 0, st, (const float *)t.H2, P + kOffW3, P + kOffB3, t.T, seed,
                       it, t.S, B, t.dY, t.loss_part);
    hipLaunchKernelGGL(k_back_hidden<kOut>, gh, blk, 0, st, (const float *)t.dY, P + kOffW3,
                       (const float *)t.Z2, B, t.dZ2);
    hipLaunchKernelGGL(k_wgrad<false>, dim3(kHid + 1, kGradChunks), blk, 0, st, (const float *)t.H2, kHid, seed, it,
                       t.S, (const float *)t.dY, kOut, B, kHid, t.gpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(((kHid + 1) * kOut + 255) / 256), blk, 0, st, (const float *)t.gpart,
                       kHid, kOut, G + kOffW3, G + kOffB3);
    hipLaunchKernelGGL(k_back_hidden<kHid>, gh, blk, 0, st, (const float *)t.dZ2, P + kOffW2,
                       (const float *)t.Z1, B, t.dZ1);
    hipLaunchKernelGGL(k_wgrad<false>, dim3(kHid + 1, kGradChunks), blk, 0, st, (const float *)t.H1, kHid, seed, it,
                       t.S, (const float *)t.dZ2, kHid, B, kHid, t.gpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(((kHid + 1) * kHid + 255) / 256), blk, 0, st, (const float *)t.gpart,
                       kHid, kHid, G + kOffW2, G + kOffB2);
    hipLaunchKernelGGL(k_wgrad<true>, dim3(kFeat + 1, kGradChunks), blk, 0, st, t.X, kFeat, seed, it, t.S,
                       (const float *)t.dZ1, kHid, B, kFeat, t.gpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(((kFeat + 1) * kHid + 255) / 256), blk, 0, st, (const float *)t.gpart,
                       kFeat, kHid, G, G + kOffB1);
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), blk, 0, st, t.P, (const float *)G, t.M, t.V,
                       kMlpParams, lr, beta1, beta2, eps, c1, c2, (const float *)t.loss_part, nparts, B,
                       t.loss_hist + it, adam);
    return hipGetLastError();
}

} // namespace nb
