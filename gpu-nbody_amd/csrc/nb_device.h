// nb_device.h -- device helpers shared by the force traversals.
#pragma once

#include "nb_internal.h"

namespace nb {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

constexpr float kFar = __builtin_huge_valf();
constexpr int kStack = 512; // >= 7 * 64 + 8 entries: the deepest possible push pattern

// v_writelane_b32: `old` with lane `lane` (wave-uniform) set to the
// wave-uniform value v; one instruction (LLVM's amdgcn.writelane intrinsic,
// which this compiler exposes under no builtin name)
__device__ int amdgcn_writelane_i32(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

__device__ __forceinline__ unsigned lane_write(unsigned old, unsigned v, int lane)
{
    return (unsigned)amdgcn_writelane_i32((int)v, lane, (int)old);
}

// The opening vote of a wave: true when every active lane's predicate holds.
// The compare result goes straight into the lane mask (no bool round trip).
__device__ __forceinline__ bool wave_all(bool pred)
{
    return __builtin_amdgcn_ballot_w64(pred) == __builtin_amdgcn_read_exec();
}

__device__ __forceinline__ void kahan_add(float &sum, float &comp, float x)
{
    const float y = x - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

__device__ __forceinline__ void flag_error(Ctl *ctl, int code)
{
    int expected = 0;
    __hip_atomic_compare_exchange_strong(&ctl->error, &expected, code, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, AGENT);
}

// Epilogue for one body: velocity correction and store (calculateforce.cl:
// 174-185), optionally the kick-drift-kick of integrate.cl:29-43 (mode 1), or
// only the acceleration into the probe buffer (mode 2).
__device__ __forceinline__ void finish_body(const Dev &d, int pos, float4 me, float fx, float fy,
                                            float fz, int mode, int step)
{
    const int b = d.sorted[pos];
    if (mode == 2) {
        d.probe[b] = make_float4(fx, fy, fz, 0.0f);
        return;
    }
    const float dt = d.dt;
    const float4 v0 = d.vel4[b];
    float vx = v0.x, vy = v0.y, vz = v0.z;
    if (step > 0) {
        const float4 a0 = d.acc4[b];
        vx = vx + ((fx - a0.x) * dt) * 0.5f;
        vy = vy + ((fy - a0.y) * dt) * 0.5f;
        vz = vz + ((fz - a0.z) * dt) * 0.5f;
    }
    d.acc4[b] = make_float4(fx, fy, fz, 0.0f);
    if (mode == 1) {
        const float dvx = (fx * dt) * 0.5f, dvy = (fy * dt) * 0.5f, dvz = (fz * dt) * 0.5f;
        const float vhx = vx + dvx, vhy = vy + dvy, vhz = vz + dvz;
        float4 p = me; // this body's (x, y, z, m)
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
        d.bnode[pos] = p; // the new record in sorted order, read by the next build
        vx = vhx + dvx;
        vy = vhy + dvy;
        vz = vhz + dvz;
    }
    d.vel4[b] = make_float4(vx, vy, vz, 0.0f);
}

} // namespace nb
