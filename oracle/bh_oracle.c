/*
 * bh_oracle.c -- CPU oracle for the Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else. The product (libnbody.so) never links, loads or calls
 * it, and no product code path falls back to it.
 *
 * It restates, sequentially and in plain C, the reference pipeline of
 * bneukom/gpu-nbody: the six kernels of kernels/nbody/<stage>.cl driven by
 * GPUBarnesHutNBodySimulation.step() (GPUBarnesHutNBodySimulation.java:249-271).
 * The node layout is the reference's: arrays of NN+1 entries, bodies in
 * [0,N), cells allocated downward from the root at index NN
 * (GPUBarnesHutNBodySimulation.java:219-227, buildtree.cl:109).
 *
 * Arithmetic choices where the reference leaves the rounding open (its
 * -cl-mad-enable contraction and driver rsqrt are unpinned) are made explicit
 * here with fmaf/fma so that the GPU kernels can reproduce them bit for bit:
 *   - squared distance:   d2 = fmaf(dx,dx, fmaf(dy,dy, fmaf(dz,dz, eps2)))
 *   - drift:              x  = fmaf(vh, dt, x)
 * Centre-of-mass accumulation runs in FP64 in child-slot order (the GPU
 * design); ORC_COM_FP32 selects the reference's FP32 arithmetic instead
 * (summarizetree.cl:107-110,161-167). Force terms are accumulated in FP64
 * with an FP64 inverse square root, so the oracle's accelerations carry no
 * FP32 summation error of their own.
 *
 * Each function cites the reference lines it follows. How every function is
 * pinned independently of this file (hand-derived trees, closed forms, brute
 * force) is listed in tests/test_oracle_pins.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXDEPTH 64 /* calculateforce.cl:12 */
#define ORC_WARPSIZE 16 /* GPUBarnesHutNBodySimulation.java:47 */

/* flags */
#define ORC_FP64_POS 1 /* leapfrog positions kept in FP64 (the GPU's precision mode) */
#define ORC_COM_FP32 2 /* reference FP32 centre-of-mass arithmetic */

/* error codes (mirrors the reference's _error word, buildtree.cl:115, calculateforce.cl:71) */
#define ORC_OK 0
#define ORC_E_POOL 1
#define ORC_E_DEPTH 2

typedef struct orc_sim {
    int n, nn;
    float theta2, eps2, dt;
    int vote, flags;
    float *x, *y, *z;    /* nn+1: bodies, then cell centres / centres of mass */
    float *vx, *vy, *vz; /* n */
    float *ax, *ay, *az; /* n */
    float *m;            /* nn+1 */
    double *x64, *y64, *z64; /* n, FP64 master positions (ORC_FP64_POS) */
    double *com;         /* 4*(nn+1): FP64 centre of mass and mass of cells */
    int *child;          /* 8*(nn+1) */
    int *count, *start, *sorted; /* nn+1 */
    int step, bottom, max_depth, error;
    float radius;
    float dq[ORC_MAXDEPTH];
    long long interactions;
    double *cq;          /* 8*(nn+1): second moments of cells (learned mode), lazily */
} orc_sim;

/* GPUBarnesHutNBodySimulation.java:219-227, with maxComputeUnits = 16 (:126). */
int orc_nodes_for(int n)
{
    int nn = n * 2;
    if (nn < 1024 * 16) nn = 1024 * 16;
    while (nn & (ORC_WARPSIZE - 1)) ++nn;
    return nn;
}

/* Opening threshold tables, calculateforce.cl:52-67: dq[0] = r^2/THETA,
 * dq[i] = dq[i-1]/4, and EPSILON added to every entry afterwards. THETA is
 * the square of the opening angle. The FP32 quotient is computed in double and
 * rounded once, which equals the correctly rounded FP32 division. */
void orc_dq_table(float radius, float theta2, float eps2, int levels, float *dq)
{
    float d0 = radius * radius;
    if (theta2 > 0.0f) d0 = (float)((double)d0 / (double)theta2);
    dq[0] = d0;
    int i;
    for (i = 1; i < levels; ++i) {
        dq[i] = 0.25f * dq[i - 1];
        dq[i - 1] += eps2;
    }
    dq[i - 1] += eps2;
}

orc_sim *orc_create(int n, float theta, float softening, float dt, int vote, int flags)
{
    if (n < 1 || vote < 1) return NULL;
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->n = n;
    s->nn = orc_nodes_for(n);
    s->theta2 = theta * theta;     /* THETA is theta squared, calculateforce.cl:15-16 */
    s->eps2 = softening * softening; /* EPSILON is the squared softening, :13-14 */
    s->dt = dt;
    s->vote = vote;
    s->flags = flags;
    size_t nodes = (size_t)s->nn + 1;
    s->x = calloc(nodes, sizeof(float));
    s->y = calloc(nodes, sizeof(float));
    s->z = calloc(nodes, sizeof(float));
    s->m = calloc(nodes, sizeof(float));
    s->vx = calloc(n, sizeof(float));
    s->vy = calloc(n, sizeof(float));
    s->vz = calloc(n, sizeof(float));
    s->ax = calloc(n, sizeof(float));
    s->ay = calloc(n, sizeof(float));
    s->az = calloc(n, sizeof(float));
    s->x64 = calloc(n, sizeof(double));
    s->y64 = calloc(n, sizeof(double));
    s->z64 = calloc(n, sizeof(double));
    s->com = calloc(4 * nodes, sizeof(double));
    s->child = malloc(8 * nodes * sizeof(int));
    for (size_t i = 0; i < 8 * nodes; ++i) s->child[i] = -1;
    s->count = calloc(nodes, sizeof(int));
    s->start = calloc(nodes, sizeof(int));
    s->sorted = calloc(nodes, sizeof(int));
    s->step = -1;     /* GPUBarnesHutNBodySimulation.java:167 */
    s->max_depth = 1; /* GPUBarnesHutNBodySimulation.java:170 */
    return s;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    free(s->cq);
    free(s->x); free(s->y); free(s->z); free(s->m);
    free(s->vx); free(s->vy); free(s->vz);
    free(s->ax); free(s->ay); free(s->az);
    free(s->x64); free(s->y64); free(s->z64);
    free(s->com); free(s->child); free(s->count); free(s->start); free(s->sorted);
    free(s);
}

/* Initial state, as UniverseGenerator.generate fills the host arrays that
 * GPUBarnesHutNBodySimulation.init copies to the device (java:134-146).
 * Accelerations start at zero (java:163-165). */
void orc_set_state(orc_sim *s, const float *x, const float *y, const float *z,
                   const float *vx, const float *vy, const float *vz, const float *m)
{
    int n = s->n;
    memcpy(s->x, x, n * sizeof(float));
    memcpy(s->y, y, n * sizeof(float));
    memcpy(s->z, z, n * sizeof(float));
    memcpy(s->vx, vx, n * sizeof(float));
    memcpy(s->vy, vy, n * sizeof(float));
    memcpy(s->vz, vz, n * sizeof(float));
    memcpy(s->m, m, n * sizeof(float));
    memset(s->ax, 0, n * sizeof(float));
    memset(s->ay, 0, n * sizeof(float));
    memset(s->az, 0, n * sizeof(float));
    for (int i = 0; i < n; ++i) {
        s->x64[i] = x[i];
        s->y64[i] = y[i];
        s->z64[i] = z[i];
    }
}

/* Overwrite accelerations and the step counter (lets a test hand the oracle a
 * GPU state in the middle of a run). */
void orc_set_accel(orc_sim *s, const float *ax, const float *ay, const float *az, int step)
{
    memcpy(s->ax, ax, s->n * sizeof(float));
    memcpy(s->ay, ay, s->n * sizeof(float));
    memcpy(s->az, az, s->n * sizeof(float));
    s->step = step;
}

void orc_set_positions64(orc_sim *s, const double *x, const double *y, const double *z)
{
    memcpy(s->x64, x, s->n * sizeof(double));
    memcpy(s->y64, y, s->n * sizeof(double));
    memcpy(s->z64, z, s->n * sizeof(double));
    for (int i = 0; i < s->n; ++i) {
        s->x[i] = (float)x[i];
        s->y[i] = (float)y[i];
        s->z[i] = (float)z[i];
    }
}

/* ---- stage 1: bounding box, boundingbox.cl:43-92 (scan), :152-160 (combine),
 *      :171-195 (root cell and counters) ---------------------------------- */
void orc_bbox_arrays(int n, const float *x, const float *y, const float *z, float *out7)
{
    float mnx = x[0], mny = y[0], mnz = z[0]; /* boundingbox.cl:49-58 */
    float mxx = x[0], mxy = y[0], mxz = z[0];
    for (int i = 0; i < n; ++i) {
        mnx = fminf(mnx, x[i]); mxx = fmaxf(mxx, x[i]);
        mny = fminf(mny, y[i]); mxy = fmaxf(mxy, y[i]);
        mnz = fminf(mnz, z[i]); mxz = fmaxf(mxz, z[i]);
    }
    out7[0] = 0.5f * (mnx + mxx); /* boundingbox.cl:171-173 */
    out7[1] = 0.5f * (mny + mxy);
    out7[2] = 0.5f * (mnz + mxz);
    out7[3] = 0.5f * fmaxf(fmaxf(mxx - mnx, mxy - mny), mxz - mnz); /* :179 */
    out7[4] = mnx; out7[5] = mny; out7[6] = mnz;
}

void orc_bbox(orc_sim *s)
{
    float r[7];
    orc_bbox_arrays(s->n, s->x, s->y, s->z, r);
    int nn = s->nn;
    s->radius = r[3];
    s->bottom = nn;          /* :180 */
    s->x[nn] = r[0];         /* :185-187 */
    s->y[nn] = r[1];
    s->z[nn] = r[2];
    s->m[nn] = -1.0f;        /* :188 */
    s->start[nn] = 0;        /* :189 */
    for (int i = 0; i < 8; ++i) s->child[8 * nn + i] = -1; /* :192-193 */
    s->step++;               /* :195 */
}

/* ---- stage 2: tree build, buildtree.cl:41-192, one body at a time -------- */
static inline int octant_of(float cx, float cy, float cz, float bx, float by, float bz)
{
    /* buildtree.cl:66-68 and :84-86: strict '<', ties go to the low side */
    int p = 0;
    if (cx < bx) p = 1;
    if (cy < by) p += 2;
    if (cz < bz) p += 4;
    return p;
}

int orc_build(orc_sim *s)
{
    const int n = s->n, nn = s->nn;
    float *X = s->x, *Y = s->y, *Z = s->z;
    int *child = s->child;
    const float radius = s->radius;
    int local_max = 1;
    for (int b = 0; b < n; ++b) {
        const float bx = X[b], by = Y[b], bz = Z[b];
        int node = nn, depth = 1; /* buildtree.cl:60-62 */
        float r = radius;
        int path = octant_of(X[nn], Y[nn], Z[nn], bx, by, bz);
        int ci = child[8 * node + path];
        while (ci >= n) { /* :77-89 */
            node = ci;
            ++depth;
            r *= 0.5f;
            path = octant_of(X[node], Y[node], Z[node], bx, by, bz);
            ci = child[8 * node + path];
        }
        if (ci == -1) { /* :98-101 */
            child[8 * node + path] = b;
        } else { /* :102-181 */
            const int locked = 8 * node + path;
            int patch = -1;
            do {
                depth++;
                const int cell = --s->bottom; /* atom_dec(_bottom) - 1, :109 */
                if (cell <= n) { /* :112-120 */
                    s->error = ORC_E_POOL;
                    s->bottom = nn;
                    return ORC_E_POOL;
                }
                if (cell > patch) patch = cell;
                float xo = (float)(path & 1) * r; /* :124-126 */
                float yo = (float)((path >> 1) & 1) * r;
                float zo = (float)((path >> 2) & 1) * r;
                r *= 0.5f; /* :128 */
                s->m[cell] = -1.0f; /* :131-132 */
                s->start[cell] = -1;
                const float cx = (X[node] - r) + xo; /* :134-136 */
                const float cy = (Y[node] - r) + yo;
                const float cz = (Z[node] - r) + zo;
                X[cell] = cx; Y[cell] = cy; Z[cell] = cz;
                for (int k = 0; k < 8; ++k) child[8 * cell + k] = -1; /* :139 */
                if (patch != cell) child[8 * node + path] = cell; /* :141-146 */
                /* re-insert the old body, :148-152 */
                child[8 * cell + octant_of(cx, cy, cz, X[ci], Y[ci], Z[ci])] = ci;
                node = cell; /* :155-161 */
                path = octant_of(cx, cy, cz, bx, by, bz);
                ci = child[8 * node + path];
            } while (ci >= 0);
            child[8 * node + path] = b; /* :169 */
            child[locked] = patch;      /* :180 */
        }
        if (depth > local_max) local_max = depth; /* :186 */
    }
    if (local_max > s->max_depth) s->max_depth = local_max; /* atom_max, :199 */
    return ORC_OK;
}

/* ---- stage 3: summarize, summarizetree.cl:55-178, cells in ascending index
 * order (children always have lower indices than their parent, since cells
 * are allocated downward), children compacted to the front (:75-81).
 *
 * FP64 mode (the GPU design): the sums run over the eight octant slots (an
 * empty slot adds 0) in the fixed pairwise order ((s0+s1)+(s2+s3))+((s4+s5)+
 * (s6+s7)); a body contributes (m, m*x, ...) from its FP32 values, a cell its
 * FP64 centre of mass. The same pass computes every cell's second moments
 * about its centre of mass for the learned mode (see orc_moments). ------- */
static double pair8(const double *a)
{
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

void orc_summarize(orc_sim *s)
{
    const int n = s->n, nn = s->nn;
    float *X = s->x, *Y = s->y, *Z = s->z, *M = s->m;
    int *child = s->child;
    const int fp32 = (s->flags & ORC_COM_FP32) != 0;
    if (!fp32 && !s->cq) s->cq = (double *)calloc(8 * ((size_t)nn + 1), sizeof(double));
    for (int node = s->bottom; node <= nn; ++node) {
        int used = 0, cnt = 0, has[8];
        double mk[8], xk[8], yk[8], zk[8], qk[8][7], bk[8];
        float m32 = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int c = child[8 * node + k];
            mk[k] = xk[k] = yk[k] = zk[k] = bk[k] = 0.0;
            for (int j = 0; j < 7; ++j) qk[k][j] = 0.0;
            has[k] = c >= 0;
            if (c < 0) continue;
            if (k != used) { /* move to front, :77-81 */
                child[8 * node + k] = -1;
                child[8 * node + used] = c;
            }
            ++used;
            if (c >= n) {
                cnt += s->count[c]; /* :98-105 with :118 */
                if (fp32) {
                    const float mc = M[c];
                    m32 += mc;
                    fx = fmaf(X[c], mc, fx); fy = fmaf(Y[c], mc, fy); fz = fmaf(Z[c], mc, fz);
                } else {
                    const double *cc = s->com + 4 * (size_t)c;
                    mk[k] = cc[3]; xk[k] = cc[0]; yk[k] = cc[1]; zk[k] = cc[2];
                    for (int j = 0; j < 7; ++j) qk[k][j] = s->cq[8 * (size_t)c + j];
                    bk[k] = s->cq[8 * (size_t)c + 7];
                }
            } else {
                cnt += 1;
                const float mc = M[c];
                if (fp32) {
                    m32 += mc;
                    fx = fmaf(X[c], mc, fx); fy = fmaf(Y[c], mc, fy); fz = fmaf(Z[c], mc, fz);
                } else {
                    mk[k] = (double)mc; xk[k] = (double)X[c]; yk[k] = (double)Y[c]; zk[k] = (double)Z[c];
                }
            }
        }
        s->count[node] = cnt; /* :160 */
        double *cc = s->com + 4 * (size_t)node;
        if (fp32) {
            /* :161-167; a cell of zero mass gets its centre at the origin
             * instead of the reference's 0*inf = NaN */
            const float inv = m32 > 0.0f ? 1.0f / m32 : 0.0f;
            X[node] = fx * inv; Y[node] = fy * inv; Z[node] = fz * inv;
            M[node] = m32;
            cc[0] = X[node]; cc[1] = Y[node]; cc[2] = Z[node]; cc[3] = m32;
        } else {
            double px[8], py[8], pz[8];
            for (int k = 0; k < 8; ++k) { px[k] = mk[k] * xk[k]; py[k] = mk[k] * yk[k]; pz[k] = mk[k] * zk[k]; }
            const double m64 = pair8(mk), sx = pair8(px), sy = pair8(py), sz = pair8(pz);
            const double inv = m64 > 0.0 ? 1.0 / m64 : 0.0;
            cc[0] = sx * inv; cc[1] = sy * inv; cc[2] = sz * inv; cc[3] = m64;
            X[node] = (float)cc[0]; Y[node] = (float)cc[1]; Z[node] = (float)cc[2];
            M[node] = (float)m64; /* the ready flag of :172, as a value */
            /* second moments: per slot, the child's own plus its mass at its
             * offset from this centre of mass (parallel-axis rule); extent:
             * the largest offset plus the child's own extent */
            double t[7][8], bmax = 0.0;
            for (int k = 0; k < 8; ++k) {
                const double dx = xk[k] - cc[0], dy = yk[k] - cc[1], dz = zk[k] - cc[2];
                if (has[k]) {
                    const double r = sqrt(fma(dx, dx, fma(dy, dy, dz * dz))) + bk[k];
                    bmax = r > bmax ? r : bmax;
                }
                const double r2 = dx * dx + dy * dy + dz * dz, mx = mk[k];
                t[0][k] = qk[k][0] + mx * (3.0 * dx * dx - r2);
                t[1][k] = qk[k][1] + mx * (3.0 * dy * dy - r2);
                t[2][k] = qk[k][2] + mx * (3.0 * dz * dz - r2);
                t[3][k] = qk[k][3] + mx * 3.0 * dx * dy;
                t[4][k] = qk[k][4] + mx * 3.0 * dx * dz;
                t[5][k] = qk[k][5] + mx * 3.0 * dy * dz;
                t[6][k] = qk[k][6] + mx * r2;
            }
            for (int j = 0; j < 7; ++j) s->cq[8 * (size_t)node + j] = pair8(t[j]);
            s->cq[8 * (size_t)node + 7] = bmax;
        }
    }
}

/* ---- stage 4: sort, sort.cl:24-75: top-down start offsets; parents have
 * higher indices than their children, so descending order sees every
 * parent before its children. ------------------------------------------- */
void orc_sort(orc_sim *s)
{
    const int n = s->n, nn = s->nn;
    for (int cell = nn; cell >= s->bottom; --cell) {
        int st = s->start[cell];
        for (int i = 0; i < 8; ++i) {
            const int ch = s->child[8 * cell + i];
            if (ch >= n) { /* :41-56 */
                s->start[ch] = st;
                st += s->count[ch];
            } else if (ch >= 0) { /* :57-64 */
                s->sorted[st] = ch;
                ++st;
            }
        }
    }
}

/* ---- stage 5: force, calculateforce.cl:52-186. Vote groups are `vote`
 * consecutive sorted positions; a cell is accepted only if every member of
 * the group accepts it (work_group_all, :145), bodies always interact. --- */
static void force_group(orc_sim *s, int g, long long *inter)
{
    const int n = s->n, nn = s->nn, W = s->vote;
    const float *X = s->x, *Y = s->y, *Z = s->z, *M = s->m;
    const int *child = s->child;
    const float eps2 = s->eps2;
    const int first = g * W;
    const int cnt = (first + W <= n) ? W : n - first;
    float px[256], py[256], pz[256];
    double ax[256], ay[256], az[256];
    int stackNode[ORC_MAXDEPTH + 1], stackPos[ORC_MAXDEPTH + 1];
    for (int l = 0; l < cnt; ++l) {
        const int b = s->sorted[first + l]; /* :101 */
        px[l] = X[b]; py[l] = Y[b]; pz[l] = Z[b];
        ax[l] = ay[l] = az[l] = 0.0;
    }
    long long ni = 0;
    int depth = 0; /* :112-116 */
    stackNode[0] = nn;
    stackPos[0] = 0;
    while (depth >= 0) { /* :122 */
        int top;
        while ((top = stackPos[depth]) < 8) { /* :125 */
            const int c = child[stackNode[depth] * 8 + top];
            stackPos[depth] = top + 1;
            if (c >= 0) {
                const float cxn = X[c], cyn = Y[c], czn = Z[c];
                int accept = 1;
                if (c >= n) { /* work_group_all(d2 >= dq[depth]), :145 */
                    for (int l = 0; l < cnt && accept; ++l) {
                        const float dx = cxn - px[l], dy = cyn - py[l], dz = czn - pz[l];
                        const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                        if (!(d2 >= s->dq[depth])) accept = 0;
                    }
                }
                if (accept) { /* :146-151 */
                    const double mc = (double)M[c];
                    for (int l = 0; l < cnt; ++l) {
                        const float dx = cxn - px[l], dy = cyn - py[l], dz = czn - pz[l];
                        const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                        const double rinv = 1.0 / sqrt((double)d2);
                        const double f = mc * rinv * rinv * rinv;
                        ax[l] += (double)dx * f;
                        ay[l] += (double)dy * f;
                        az[l] += (double)dz * f;
                    }
                    ni += cnt;
                } else { /* push, :155-160 */
                    depth++;
                    stackNode[depth] = c;
                    stackPos[depth] = 0;
                }
            } else {
                depth = depth - 1 > 0 ? depth - 1 : 0; /* :165-167 */
            }
        }
        depth--; /* :171 */
    }
    const float dt = s->dt;
    for (int l = 0; l < cnt; ++l) {
        const int b = s->sorted[first + l];
        const float fx = (float)ax[l], fy = (float)ay[l], fz = (float)az[l];
        if (s->step > 0) { /* :174-179 */
            s->vx[b] = s->vx[b] + ((fx - s->ax[b]) * dt) * 0.5f;
            s->vy[b] = s->vy[b] + ((fy - s->ay[b]) * dt) * 0.5f;
            s->vz[b] = s->vz[b] + ((fz - s->az[b]) * dt) * 0.5f;
        }
        s->ax[b] = fx; /* :183-185 */
        s->ay[b] = fy;
        s->az[b] = fz;
    }
    *inter += ni;
}

int orc_force(orc_sim *s)
{
    orc_dq_table(s->radius, s->theta2, s->eps2, s->max_depth, s->dq); /* :52-67 */
    if (s->max_depth > ORC_MAXDEPTH) { /* :69-73 */
        s->error = ORC_E_DEPTH;
        return ORC_E_DEPTH;
    }
    if (s->vote > 256) return -1;
    const int groups = (s->n + s->vote - 1) / s->vote;
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
    for (int g = 0; g < groups; ++g) force_group(s, g, &total);
    s->interactions = total;
    return ORC_OK;
}

/* Force on a subset of vote groups only (for sampled checks at large N).
 * Accelerations of bodies outside the listed groups are left untouched. */
int orc_force_groups(orc_sim *s, const int *groups, int ngroups)
{
    orc_dq_table(s->radius, s->theta2, s->eps2, s->max_depth, s->dq);
    if (s->max_depth > ORC_MAXDEPTH) {
        s->error = ORC_E_DEPTH;
        return ORC_E_DEPTH;
    }
    if (s->vote > 256) return -1;
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int i = 0; i < ngroups; ++i) force_group(s, groups[i], &total);
    s->interactions = total;
    return ORC_OK;
}

/* The CPU baseline's FP32 force pass over the listed vote groups (cpu_fp32.c:
 * the reference's FP32 arithmetic, calculateforce.cl:143-151, vectorized over
 * a group's lanes). Accelerations by sorted position into ax/ay/az (n each);
 * the oracle's own accelerations are not touched. Returns the interactions. */
long long cpu_force_fp32(int n, int nn, const float *X, const float *Y, const float *Z, const float *M,
                         const int *child, const int *sorted, const float *dq, float eps2, int vote,
                         const int *groups, int ngroups, float *ax_out, float *ay_out, float *az_out);

long long orc_force_fp32_groups(orc_sim *s, const int *groups, int ngroups, float *ax, float *ay, float *az)
{
    orc_dq_table(s->radius, s->theta2, s->eps2, s->max_depth, s->dq);
    if (s->max_depth > ORC_MAXDEPTH) return -1;
    return cpu_force_fp32(s->n, s->nn, s->x, s->y, s->z, s->m, s->child, s->sorted, s->dq, s->eps2, s->vote,
                          groups, ngroups, ax, ay, az);
}

/* ---- stage 6: integrate, integrate.cl:27-44 (kick, drift, kick with the
 * acceleration just computed). ------------------------------------------ */
void orc_integrate(orc_sim *s)
{
    const float dt = s->dt;
    const int fp64 = (s->flags & ORC_FP64_POS) != 0;
    for (int i = 0; i < s->n; ++i) {
        const float dvx = (s->ax[i] * dt) * 0.5f; /* :29-31 */
        const float dvy = (s->ay[i] * dt) * 0.5f;
        const float dvz = (s->az[i] * dt) * 0.5f;
        const float vhx = s->vx[i] + dvx; /* :33-35 */
        const float vhy = s->vy[i] + dvy;
        const float vhz = s->vz[i] + dvz;
        if (fp64) {
            s->x64[i] = fma((double)vhx, (double)dt, s->x64[i]);
            s->y64[i] = fma((double)vhy, (double)dt, s->y64[i]);
            s->z64[i] = fma((double)vhz, (double)dt, s->z64[i]);
            s->x[i] = (float)s->x64[i];
            s->y[i] = (float)s->y64[i];
            s->z[i] = (float)s->z64[i];
        } else {
            s->x[i] = fmaf(vhx, dt, s->x[i]); /* :37-39 */
            s->y[i] = fmaf(vhy, dt, s->y[i]);
            s->z[i] = fmaf(vhz, dt, s->z[i]);
        }
        s->vx[i] = vhx + dvx; /* :41-43 */
        s->vy[i] = vhy + dvy;
        s->vz[i] = vhz + dvz;
    }
}

/* ---- step(): GPUBarnesHutNBodySimulation.java:249-271 --------------------- */
int orc_step(orc_sim *s, int nsteps)
{
    for (int k = 0; k < nsteps; ++k) {
        orc_bbox(s);
        int e = orc_build(s);
        if (e) return e;
        orc_summarize(s);
        orc_sort(s);
        e = orc_force(s);
        if (e) return e;
        orc_integrate(s);
    }
    return ORC_OK;
}

/* Run only stages [first, last] of one step (1 = bbox ... 6 = integrate),
 * the way GPUBarnesHutNBodySimulation.executeSimulationKernel (java:273-291)
 * lets a caller drive them one by one. */
int orc_run_stages(orc_sim *s, int first, int last)
{
    for (int st = first; st <= last; ++st) {
        int e = 0;
        switch (st) {
        case 1: orc_bbox(s); break;
        case 2: e = orc_build(s); break;
        case 3: orc_summarize(s); break;
        case 4: orc_sort(s); break;
        case 5: e = orc_force(s); break;
        case 6: orc_integrate(s); break;
        default: return -1;
        }
        if (e) return e;
    }
    return ORC_OK;
}

/* ---- read-back helpers ---------------------------------------------------- */
void orc_get_state(const orc_sim *s, float *x, float *y, float *z, float *vx, float *vy, float *vz)
{
    const int n = s->n;
    memcpy(x, s->x, n * sizeof(float));
    memcpy(y, s->y, n * sizeof(float));
    memcpy(z, s->z, n * sizeof(float));
    memcpy(vx, s->vx, n * sizeof(float));
    memcpy(vy, s->vy, n * sizeof(float));
    memcpy(vz, s->vz, n * sizeof(float));
}

void orc_get_accel(const orc_sim *s, float *ax, float *ay, float *az)
{
    memcpy(ax, s->ax, s->n * sizeof(float));
    memcpy(ay, s->ay, s->n * sizeof(float));
    memcpy(az, s->az, s->n * sizeof(float));
}

void orc_get_positions64(const orc_sim *s, double *x, double *y, double *z)
{
    memcpy(x, s->x64, s->n * sizeof(double));
    memcpy(y, s->y64, s->n * sizeof(double));
    memcpy(z, s->z64, s->n * sizeof(double));
}

void orc_get_sorted(const orc_sim *s, int *sorted)
{
    memcpy(sorted, s->sorted, s->n * sizeof(int));
}

/* scalars: [0]=step [1]=bottom [2]=max_depth [3]=error [4]=cells [5]=nn */
void orc_get_scalars(const orc_sim *s, long long *out)
{
    out[0] = s->step;
    out[1] = s->bottom;
    out[2] = s->max_depth;
    out[3] = s->error;
    out[4] = s->nn - s->bottom + 1;
    out[5] = s->nn;
    out[6] = s->interactions;
}

/* root centre and radius after stage 1 */
void orc_get_root(const orc_sim *s, float *out4)
{
    out4[0] = s->x[s->nn];
    out4[1] = s->y[s->nn];
    out4[2] = s->z[s->nn];
    out4[3] = s->radius;
}

/* Leaf membership: for every body its octant path from the root (3 bits per
 * level, level 0 in the lowest bits of paths[3*i], 21 levels per word) and its
 * depth (number of octants on the path). Cell numbering is timing-dependent in
 * the reference; the path is not. The slot index is the octant only before
 * summarize compacts the child slots, so paths are read after stage 2.
 * Returns the number of bodies reached, or -1 on a tree deeper than 63. */
static void put_path(unsigned long long *dst, const unsigned long long *cur, int levels)
{
    for (int w = 0; w < 3; ++w) {
        const int lv = levels - 21 * w;
        if (lv <= 0) dst[w] = 0;
        else if (lv >= 21) dst[w] = cur[w] & ((1ULL << 63) - 1);
        else dst[w] = cur[w] & ((1ULL << (3 * lv)) - 1);
    }
}

int orc_get_paths(const orc_sim *s, unsigned long long *paths, int *depth)
{
    const int n = s->n, nn = s->nn;
    int reached = 0;
    int stackNode[ORC_MAXDEPTH + 2], stackPos[ORC_MAXDEPTH + 2];
    unsigned long long cur[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) depth[i] = 0;
    int d = 0;
    stackNode[0] = nn;
    stackPos[0] = 0;
    while (d >= 0) {
        if (stackPos[d] >= 8) { --d; continue; }
        const int k = stackPos[d]++;
        const int c = s->child[8 * stackNode[d] + k];
        if (c < 0) continue;
        const int w = d / 21, sh = 3 * (d % 21);
        cur[w] = (cur[w] & ~(7ULL << sh)) | ((unsigned long long)k << sh);
        if (c < n) {
            put_path(paths + 3 * (size_t)c, cur, d + 1);
            depth[c] = d + 1;
            ++reached;
        } else {
            if (d + 1 >= 63) return -1;
            ++d;
            stackNode[d] = c;
            stackPos[d] = 0;
        }
    }
    return reached;
}

/* Per-cell data in the reference layout for [bottom, nn]: count, FP32 centre
 * of mass and mass. out must hold (nn-bottom+1) entries each. */
void orc_get_cells(const orc_sim *s, int *count, float *x, float *y, float *z, float *m, int *child8)
{
    const int b = s->bottom, nn = s->nn;
    for (int c = b; c <= nn; ++c) {
        count[c - b] = s->count[c];
        x[c - b] = s->x[c];
        y[c - b] = s->y[c];
        z[c - b] = s->z[c];
        m[c - b] = s->m[c];
        for (int k = 0; k < 8; ++k) child8[8 * (c - b) + k] = s->child[8 * c + k];
    }
}

/* ---- FP64 direct sum on sampled targets (the error yardstick) ------------ */
void orc_direct_fp64(int n, const float *x, const float *y, const float *z, const float *m,
                     double eps2, int ns, const int *idx, double *out_ax, double *out_ay, double *out_az)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < ns; ++t) {
        const int i = idx[t];
        const double xi = x[i], yi = y[i], zi = z[i];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = 0; j < n; ++j) {
            const double dx = (double)x[j] - xi, dy = (double)y[j] - yi, dz = (double)z[j] - zi;
            const double d2 = dx * dx + dy * dy + dz * dz + eps2;
            const double r = 1.0 / sqrt(d2);
            const double f = (double)m[j] * r * r * r;
            sx += dx * f;
            sy += dy * f;
            sz += dz * f;
        }
        out_ax[t] = sx;
        out_ay[t] = sy;
        out_az[t] = sz;
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int t)
{
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* ---- MLP of the learned correction (no reference counterpart) ------------
 * 24 -> 64 -> 64 -> 3, GELU (tanh form) on both hidden layers. Parameter
 * layout (5955 values): W1[24][64], b1[64], W2[64][64], b2[64], W3[64][3],
 * b3[3], W[i][j] connecting input i to output j. FP64 throughout: this is the
 * yardstick for the engine's FP32 training kernels and BF16 inference. */
#define ORC_MLP_IN 24
#define ORC_MLP_H 64
#define ORC_MLP_OUT 3
#define ORC_MLP_PARAMS (ORC_MLP_IN * ORC_MLP_H + ORC_MLP_H + ORC_MLP_H * ORC_MLP_H + ORC_MLP_H + ORC_MLP_H * ORC_MLP_OUT + ORC_MLP_OUT)

static double gelu(double x)
{
    const double c = 0.7978845608028654; /* sqrt(2/pi) */
    return 0.5 * x * (1.0 + tanh(c * (x + 0.044715 * x * x * x)));
}

static double gelu_grad(double x)
{
    const double c = 0.7978845608028654;
    const double u = c * (x + 0.044715 * x * x * x);
    const double t = tanh(u);
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * c * (1.0 + 3.0 * 0.044715 * x * x);
}

int orc_mlp_params(void) { return ORC_MLP_PARAMS; }

static void mlp_one(const double *P, const float *x, double *z1, double *h1, double *z2, double *h2, double *y)
{
    const double *W1 = P, *b1 = W1 + ORC_MLP_IN * ORC_MLP_H, *W2 = b1 + ORC_MLP_H;
    const double *b2 = W2 + ORC_MLP_H * ORC_MLP_H, *W3 = b2 + ORC_MLP_H, *b3 = W3 + ORC_MLP_H * ORC_MLP_OUT;
    for (int j = 0; j < ORC_MLP_H; ++j) {
        double s = b1[j];
        for (int i = 0; i < ORC_MLP_IN; ++i) s += (double)x[i] * W1[i * ORC_MLP_H + j];
        z1[j] = s;
        h1[j] = gelu(s);
    }
    for (int j = 0; j < ORC_MLP_H; ++j) {
        double s = b2[j];
        for (int i = 0; i < ORC_MLP_H; ++i) s += h1[i] * W2[i * ORC_MLP_H + j];
        z2[j] = s;
        h2[j] = gelu(s);
    }
    for (int j = 0; j < ORC_MLP_OUT; ++j) {
        double s = b3[j];
        for (int i = 0; i < ORC_MLP_H; ++i) s += h2[i] * W3[i * ORC_MLP_OUT + j];
        y[j] = s;
    }
}

/* y[n][3] = MLP(x[n][24]) */
void orc_mlp_forward(const double *P, const float *x, int n, double *y)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        double z1[ORC_MLP_H], h1[ORC_MLP_H], z2[ORC_MLP_H], h2[ORC_MLP_H];
        mlp_one(P, x + (size_t)s * ORC_MLP_IN, z1, h1, z2, h2, y + (size_t)s * ORC_MLP_OUT);
    }
}

/* Mean squared error over n samples and its gradient with respect to every
 * parameter (backpropagation), FP64. Returns the loss. */
double orc_mlp_loss_grad(const double *P, const float *x, const float *t, int n, double *grad)
{
    const double *W2 = P + ORC_MLP_IN * ORC_MLP_H + ORC_MLP_H;
    const double *W3 = W2 + ORC_MLP_H * ORC_MLP_H + ORC_MLP_H;
    for (int k = 0; k < ORC_MLP_PARAMS; ++k) grad[k] = 0.0;
    double *gW1 = grad, *gb1 = gW1 + ORC_MLP_IN * ORC_MLP_H, *gW2 = gb1 + ORC_MLP_H;
    double *gb2 = gW2 + ORC_MLP_H * ORC_MLP_H, *gW3 = gb2 + ORC_MLP_H, *gb3 = gW3 + ORC_MLP_H * ORC_MLP_OUT;
    double loss = 0.0;
    for (int s = 0; s < n; ++s) {
        const float *xs = x + (size_t)s * ORC_MLP_IN;
        double z1[ORC_MLP_H], h1[ORC_MLP_H], z2[ORC_MLP_H], h2[ORC_MLP_H], y[ORC_MLP_OUT];
        mlp_one(P, xs, z1, h1, z2, h2, y);
        double dy[ORC_MLP_OUT], dz2[ORC_MLP_H], dz1[ORC_MLP_H];
        for (int j = 0; j < ORC_MLP_OUT; ++j) {
            const double e = y[j] - (double)t[(size_t)s * ORC_MLP_OUT + j];
            loss += e * e;
            dy[j] = 2.0 * e / ((double)n * ORC_MLP_OUT);
            gb3[j] += dy[j];
        }
        for (int i = 0; i < ORC_MLP_H; ++i) {
            double dh = 0.0;
            for (int j = 0; j < ORC_MLP_OUT; ++j) {
                gW3[i * ORC_MLP_OUT + j] += h2[i] * dy[j];
                dh += W3[i * ORC_MLP_OUT + j] * dy[j];
            }
            dz2[i] = dh * gelu_grad(z2[i]);
            gb2[i] += dz2[i];
        }
        for (int i = 0; i < ORC_MLP_H; ++i) {
            double dh = 0.0;
            for (int j = 0; j < ORC_MLP_H; ++j) {
                gW2[i * ORC_MLP_H + j] += h1[i] * dz2[j];
                dh += W2[i * ORC_MLP_H + j] * dz2[j];
            }
            dz1[i] = dh * gelu_grad(z1[i]);
            gb1[i] += dz1[i];
        }
        for (int i = 0; i < ORC_MLP_IN; ++i)
            for (int j = 0; j < ORC_MLP_H; ++j) gW1[i * ORC_MLP_H + j] += (double)xs[i] * dz1[j];
    }
    return loss / ((double)n * ORC_MLP_OUT);
}

/* ---- coarse traversal of the learned correction (no reference counterpart)
 *
 * Second moments (computed by orc_summarize in FP64 mode): every cell's
 * traceless Q = sum m (3 s s^T - |s|^2 I) (xx, yy, zz, xy, xz, yz) and trace
 * T = sum m |s|^2 about its FP64 centre of mass, moved to the parent by the
 * parallel-axis rule, summed over the octant slots in the pairwise order of
 * the centre of mass; and its extent b = max over children of
 * (|child centre - centre| + child extent), an upper bound of the distance of
 * any of its bodies from its centre of mass (bodies have extent 0). Bodies are
 * point masses at their FP32 positions.
 *
 * orc_coarse: the wave-uniform traversal of the force stage, restated per vote
 * group: accepted cells contribute
 *     m d / s^3 - Q d / s^5 + 5/2 (d.Q.d - T eps^2) d / s^7
 * (d = centre of mass - body, s^2 = |d|^2 + eps^2, the second-order expansion
 * of the softened kernel), bodies m d / s^3. A cell is accepted when every
 * lane of the group accepts it. With mac = ORC_MAC_GEOMETRIC a lane accepts
 * when s^2 >= dq[depth] (the opening table of `theta`, as in the force stage).
 * With mac = ORC_MAC_EXTENT a lane accepts when both
 *     s^2 >= b^2 / theta^2 + eps^2        (the extent test: |d| >= b / theta)
 *     s^2 >= (m b^3)^(2/5) (alpha^2 |a_old|^2)^(-1/5)
 *                                          (m b^3 / s^5, the size of the first
 *                                           term left out, within alpha |a_old|)
 * hold, with a_old the body's acceleration of the previous step (the test is
 * skipped for a lane while step <= 0, when the force stage has no previous
 * acceleration either (calculateforce.cl:174-179), when |a_old| = 0, or when
 * alpha <= 0). Opening decisions are made in
 * FP32 with the engine's operation order: s^2 = fmaf chain, thr = fmaf(b^2,
 * 1/theta^2, eps^2), the relative threshold the FP32 product of the cell's
 * factor (m b^3)^(2/5) (FP64 exp2(0.4 log2(m b^3)), rounded) and the body's factor
 * (alpha^2 |a_old|^2)^(-1/5) (FP64 pow of the FP32 alpha^2 |a_old|^2,
 * rounded; 0 while the test is off), with b^2 rounded from FP64 per child;
 * the force terms are summed in FP64. For each
 * body of the listed groups it writes the acceleration (acc, 3 per body) and,
 * when feat is not NULL, the 24 features of the engine's coarse kernel and the
 * frame vector e2 (3 per body). Output row of lane l of the i-th listed group:
 * i * vote + l. */
void orc_moments(orc_sim *s)
{
    /* computed by orc_summarize (FP64 mode); kept for callers that ask */
    (void)s;
}

/* second moments of a cell (for tests): out8 = Qxx, Qyy, Qzz, Qxy, Qxz, Qyz, T, extent */
int orc_get_moments(const orc_sim *s, int cell, double *out8)
{
    if (!s->cq || cell < s->n || cell > s->nn) return -1;
    for (int j = 0; j < 8; ++j) out8[j] = s->cq[8 * (size_t)cell + j];
    return 0;
}

#define ORC_MAC_GEOMETRIC 0
#define ORC_MAC_EXTENT 1

typedef struct coarse_cfg {
    int mac;
    const float *dq; /* geometric: the opening table */
    float it2;       /* extent: 1 / theta^2 */
    float alpha2;    /* extent: alpha^2 (<= 0: no relative test) */
} coarse_cfg;

/* per child: the threshold fmaf(b^2, 1/theta^2, eps^2) of b^2 rounded to FP32
 * from the FP64 extent, and the relative test's cell factor (m b^3)^(2/5)
 * rounded from FP64 */
static void extent_terms(const orc_sim *s, int c, float it2, float *thr, float *mb3f)
{
    const double b = s->cq[8 * (size_t)c + 7], mc = s->com[4 * (size_t)c + 3];
    const double mb3 = mc * ((b * b) * b);
    *thr = fmaf((float)(b * b), it2, s->eps2);
    *mb3f = (float)exp2(0.4 * log2(mb3));
}

static void coarse_group(const orc_sim *s, const coarse_cfg *cfg, int g, int slot0, float *feat, double *acc,
                         double *e2out, long long *inter)
{
    const int n = s->n, nn = s->nn, W = s->vote;
    const float *X = s->x, *Y = s->y, *Z = s->z, *M = s->m;
    const int *child = s->child;
    const float eps2 = s->eps2;
    const int first = g * W;
    const int cnt = (first + W <= n) ? W : n - first;
    if (cnt <= 0) return;
    float px[256], py[256], pz[256], a5[256];
    /* per lane: total, bodies, quadrupole part */
    double A[256][3], Bd[256][3], Qp[256][3];
    double smc = 0.0, smb = 0.0;
    long long nc = 0, nbod = 0;
    int stackNode[ORC_MAXDEPTH + 1], stackPos[ORC_MAXDEPTH + 1];
    for (int l = 0; l < cnt; ++l) {
        const int b = s->sorted[first + l];
        px[l] = X[b]; py[l] = Y[b]; pz[l] = Z[b];
        for (int j = 0; j < 3; ++j) A[l][j] = Bd[l][j] = Qp[l][j] = 0.0;
        const float a2 = fmaf(s->ax[b], s->ax[b], fmaf(s->ay[b], s->ay[b], s->az[b] * s->az[b]));
        a5[l] = (cfg->alpha2 > 0.0f && s->step > 0 && a2 > 0.0f) ? (float)pow((double)(cfg->alpha2 * a2), -0.2) : 0.0f;
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
            const float cxn = X[c], cyn = Y[c], czn = Z[c];
            int accept = 1;
            float thr = 0.0f;
            if (c >= n) {
                float mb3f = 0.0f;
                if (cfg->mac == ORC_MAC_EXTENT) {
                    extent_terms(s, c, cfg->it2, &thr, &mb3f);
                } else {
                    thr = cfg->dq[depth];
                }
                for (int l = 0; l < cnt && accept; ++l) {
                    const float dx = cxn - px[l], dy = cyn - py[l], dz = czn - pz[l];
                    const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                    if (!(d2 >= thr)) accept = 0;
                    if (accept && cfg->mac == ORC_MAC_EXTENT && !(mb3f * a5[l] <= d2)) accept = 0;
                }
            }
            if (!accept) {
                depth++;
                stackNode[depth] = c;
                stackPos[depth] = 0;
                continue;
            }
            const double mc = (double)M[c];
            const double *q = c >= n ? s->cq + 8 * (size_t)c : NULL;
            if (c >= n) { smc += mc; ++nc; } else { smb += mc; ++nbod; }
            for (int l = 0; l < cnt; ++l) {
                const float fdx = cxn - px[l], fdy = cyn - py[l], fdz = czn - pz[l];
                const float d2f = fmaf(fdx, fdx, fmaf(fdy, fdy, fmaf(fdz, fdz, eps2)));
                const double dx = fdx, dy = fdy, dz = fdz, d2 = d2f;
                const double r2i = 1.0 / d2, ri = sqrt(r2i), r3 = r2i * ri;
                const double f = mc * r3;
                A[l][0] += dx * f; A[l][1] += dy * f; A[l][2] += dz * f;
                if (!q) {
                    Bd[l][0] += dx * f; Bd[l][1] += dy * f; Bd[l][2] += dz * f;
                    continue;
                }
                const double Qx = q[0] * dx + q[3] * dy + q[4] * dz;
                const double Qy = q[3] * dx + q[1] * dy + q[5] * dz;
                const double Qz = q[4] * dx + q[5] * dy + q[2] * dz;
                const double dQd = Qx * dx + Qy * dy + Qz * dz;
                const double r5 = r3 * r2i;
                const double c2 = 2.5 * (dQd - q[6] * (double)eps2) * r5 * r2i;
                const double gx = c2 * dx - r5 * Qx, gy = c2 * dy - r5 * Qy, gz = c2 * dz - r5 * Qz;
                A[l][0] += gx; A[l][1] += gy; A[l][2] += gz;
                Qp[l][0] += gx; Qp[l][1] += gy; Qp[l][2] += gz;
            }
            ni += cnt;
        }
        depth--;
    }
    *inter += ni;
    /* group centroid and radius, mean acceleration, mean body and quadrupole parts */
    double gx = 0, gy = 0, gz = 0, rg = 0, am[3] = {0, 0, 0}, bm[3] = {0, 0, 0}, qm[3] = {0, 0, 0};
    for (int l = 0; l < cnt; ++l) {
        gx += px[l]; gy += py[l]; gz += pz[l];
        for (int j = 0; j < 3; ++j) { am[j] += A[l][j]; bm[j] += Bd[l][j]; qm[j] += Qp[l][j]; }
    }
    gx /= cnt; gy /= cnt; gz /= cnt;
    for (int j = 0; j < 3; ++j) { am[j] /= cnt; bm[j] /= cnt; qm[j] /= cnt; }
    for (int l = 0; l < cnt; ++l) {
        const double ox = px[l] - gx, oy = py[l] - gy, oz = pz[l] - gz;
        const double r = sqrt(ox * ox + oy * oy + oz * oz);
        if (r > rg) rg = r;
    }
    for (int l = 0; l < cnt; ++l) {
        const size_t o = (size_t)slot0 + l;
        double *a = acc + 3 * o;
        a[0] = A[l][0]; a[1] = A[l][1]; a[2] = A[l][2];
        if (!feat) continue;
        float *f = feat + ORC_MLP_IN * o;
        for (int i = 0; i < ORC_MLP_IN; ++i) f[i] = 0.0f;
        double e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0}, e3[3];
        const double An = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (An > 0.0) {
            const double inv = 1.0 / An;
            for (int j = 0; j < 3; ++j) e1[j] = a[j] * inv;
            const double *qv = Qp[l];
            const double qe1 = qv[0] * e1[0] + qv[1] * e1[1] + qv[2] * e1[2];
            double w[3] = {qv[0] - qe1 * e1[0], qv[1] - qe1 * e1[1], qv[2] - qe1 * e1[2]};
            double Wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            if (!(Wn > 0.0)) {
                /* the coordinate axis least aligned with e1 */
                const double fx = fabs(e1[0]), fy = fabs(e1[1]), fz = fabs(e1[2]);
                double t[3] = {0, 0, 0};
                if (fx <= fy && fx <= fz) t[0] = 1; else if (fy <= fz) t[1] = 1; else t[2] = 1;
                const double te1 = t[0] * e1[0] + t[1] * e1[1] + t[2] * e1[2];
                for (int j = 0; j < 3; ++j) w[j] = t[j] - te1 * e1[j];
                Wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            }
            for (int j = 0; j < 3; ++j) e2[j] = w[j] / Wn;
            e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
            e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
            e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
#define ORC_LOC(V, I, S)                                                                           \
    do {                                                                                           \
        f[I] = (float)((V[0] * e1[0] + V[1] * e1[1] + V[2] * e1[2]) * (S));                        \
        f[(I) + 1] = (float)((V[0] * e2[0] + V[1] * e2[1] + V[2] * e2[2]) * (S));                  \
        f[(I) + 2] = (float)((V[0] * e3[0] + V[1] * e3[1] + V[2] * e3[2]) * (S));                  \
    } while (0)
            ORC_LOC(Bd[l], 0, inv);
            f[3] = (float)(qe1 * inv);
            f[4] = (float)((qv[0] * e2[0] + qv[1] * e2[1] + qv[2] * e2[2]) * inv);
            const double tv[3] = {a[0] - am[0], a[1] - am[1], a[2] - am[2]};
            ORC_LOC(tv, 5, inv);
            if (rg > 0.0) {
                const double o3[3] = {px[l] - gx, py[l] - gy, pz[l] - gz};
                ORC_LOC(o3, 8, 1.0 / rg);
            }
            ORC_LOC(bm, 11, inv);
            ORC_LOC(qm, 14, inv);
            f[17] = (float)log1p((double)nc);
            f[18] = (float)log1p((double)nbod);
            const double qn = sqrt(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2]);
            const double bn = sqrt(Bd[l][0] * Bd[l][0] + Bd[l][1] * Bd[l][1] + Bd[l][2] * Bd[l][2]);
            f[19] = (float)log(fmax(qn * inv, 1e-30));
            f[20] = (float)log(fmax(bn * inv, 1e-30));
            f[21] = (smc + smb) > 0.0 ? (float)(smc / (smc + smb)) : 0.0f;
            f[22] = (float)log(fmax(rg / (double)s->radius, 1e-30));
            const double mv[3] = {a[0] - Bd[l][0] - qv[0], a[1] - Bd[l][1] - qv[1], a[2] - Bd[l][2] - qv[2]};
            const double mn = sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
            f[23] = (mn > 0.0 && bn > 0.0) ? (float)((mv[0] * Bd[l][0] + mv[1] * Bd[l][1] + mv[2] * Bd[l][2]) / (mn * bn)) : 0.0f;
#undef ORC_LOC
        }
        if (e2out) {
            e2out[3 * o] = e2[0]; e2out[3 * o + 1] = e2[1]; e2out[3 * o + 2] = e2[2];
        }
    }
}

long long orc_coarse(orc_sim *s, int mac, float theta, float alpha, const int *groups, int ngroups, float *feat,
                     double *acc, double *e2)
{
    float dq[ORC_MAXDEPTH];
    if (!s->cq || s->vote > 256 || s->max_depth > ORC_MAXDEPTH) return -1;
    if (mac != ORC_MAC_GEOMETRIC && mac != ORC_MAC_EXTENT) return -1;
    coarse_cfg cfg;
    const float theta2 = theta * theta;
    cfg.mac = mac;
    cfg.dq = dq;
    cfg.it2 = 1.0f / theta2;
    cfg.alpha2 = alpha > 0.0f ? (float)((double)alpha * (double)alpha) : 0.0f;
    orc_dq_table(s->radius, theta2, s->eps2, s->max_depth, dq);
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int i = 0; i < ngroups; ++i) coarse_group(s, &cfg, groups[i], i * s->vote, feat, acc, e2, &total);
    return total;
}
