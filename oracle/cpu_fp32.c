/*
 * cpu_fp32.c -- the CPU baseline's force pass: the reference force kernel's
 * arithmetic (calculateforce.cl:99-186) in FP32 on the host cores, for timing
 * beside the GPU. Test and benchmark infrastructure, like the rest of
 * oracle/: nothing in the engine calls it.
 *
 * Differences from the oracle's own force pass (bh_oracle.c, force_group):
 * FP32 accumulation and an FP32 1/sqrt, as the reference computes
 * (calculateforce.cl:143-151; distSquared = dx dx + dy dy + dz dz + EPSILON,
 * f = m r^-3, acc += d f), where the oracle accumulates in FP64 to serve as
 * the parity reference. The traversal is the reference's: one explicit stack
 * per vote group of `vote` consecutive sorted bodies, bodies always
 * interacted, a cell taken only when every member of the group accepts it
 * (work_group_all, :145), opened otherwise (:155-160). The group's lanes are
 * the SIMD lanes of the host (16 FP32 lanes of AVX-512 at the reference's
 * vote width of 16), and groups are spread over the OpenMP threads.
 */
#include <math.h>
#include <stddef.h>

#define CPU_MAXDEPTH 64
#define CPU_MAXVOTE 64

/* One vote group; returns its interactions. Lanes past the last body sit
 * far away (they accept every cell and their sums are dropped). */
__attribute__((target_clones("avx512f", "avx2", "default")))
static long long group_fp32(int n, int nn, const float *X, const float *Y, const float *Z, const float *M,
                            const int *child, const int *sorted, const float *dq, float eps2, int vote, int g,
                            float *ax_out, float *ay_out, float *az_out)
{
    float px[CPU_MAXVOTE], py[CPU_MAXVOTE], pz[CPU_MAXVOTE];
    float ax[CPU_MAXVOTE], ay[CPU_MAXVOTE], az[CPU_MAXVOTE];
    int stackNode[CPU_MAXDEPTH + 1], stackPos[CPU_MAXDEPTH + 1];
    const int first = g * vote;
    const int cnt = (first + vote <= n) ? vote : n - first;
    for (int l = 0; l < vote; ++l) {
        const int b = l < cnt ? sorted[first + l] : -1;
        px[l] = b >= 0 ? X[b] : 1e30f;
        py[l] = b >= 0 ? Y[b] : 1e30f;
        pz[l] = b >= 0 ? Z[b] : 1e30f;
        ax[l] = ay[l] = az[l] = 0.0f;
    }
    long long ni = 0;
    int depth = 0;
    stackNode[0] = nn;
    stackPos[0] = 0;
    while (depth >= 0) {
        int top;
        while ((top = stackPos[depth]) < 8) {
            const int c = child[stackNode[depth] * 8 + top];
            stackPos[depth] = top + 1;
            if (c < 0) {
                depth = depth - 1 > 0 ? depth - 1 : 0;
                continue;
            }
            const float cx = X[c], cy = Y[c], cz = Z[c], mc = M[c];
            int accept = 1;
            if (c >= n) {
                const float thr = dq[depth];
                int all = 1;
#pragma omp simd reduction(& : all)
                for (int l = 0; l < vote; ++l) {
                    const float dx = cx - px[l], dy = cy - py[l], dz = cz - pz[l];
                    const float d2 = dx * dx + dy * dy + dz * dz + eps2;
                    all &= d2 >= thr;
                }
                accept = all;
            }
            if (accept) {
#pragma omp simd
                for (int l = 0; l < vote; ++l) {
                    const float dx = cx - px[l], dy = cy - py[l], dz = cz - pz[l];
                    const float d2 = dx * dx + dy * dy + dz * dz + eps2;
                    const float r = 1.0f / sqrtf(d2);
                    const float f = mc * r * r * r;
                    ax[l] += dx * f;
                    ay[l] += dy * f;
                    az[l] += dz * f;
                }
                ni += cnt;
            } else if (depth + 1 <= CPU_MAXDEPTH) {
                ++depth;
                stackNode[depth] = c;
                stackPos[depth] = 0;
            }
        }
        --depth;
    }
    for (int l = 0; l < cnt; ++l) {
        ax_out[first + l] = ax[l];
        ay_out[first + l] = ay[l];
        az_out[first + l] = az[l];
    }
    return ni;
}

/* Accelerations of the listed vote groups into a*_out (by sorted position);
 * returns the interactions, or -1 for a vote width above 64. */
long long cpu_force_fp32(int n, int nn, const float *X, const float *Y, const float *Z, const float *M,
                         const int *child, const int *sorted, const float *dq, float eps2, int vote,
                         const int *groups, int ngroups, float *ax_out, float *ay_out, float *az_out)
{
    if (vote < 1 || vote > CPU_MAXVOTE) return -1;
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int i = 0; i < ngroups; ++i)
        total += group_fp32(n, nn, X, Y, Z, M, child, sorted, dq, eps2, vote, groups[i], ax_out, ay_out, az_out);
    return total;
}
