/*
 * nbody.h -- C ABI of the MI355X Barnes-Hut engine (libnbody.so).
 *
 * A drop-in for the kernels/nbody pipeline of bneukom/gpu-nbody. The entry
 * points replace what the reference's host class exposes through its plugin
 * API (src/ch/fhnw/woipv/nbody/simulation/AbstractNBodySimulation.java:5-44)
 * and what GPUBarnesHutNBodySimulation does behind it; each declaration names
 * the reference interface it replaces. Plain C types only: an opaque handle,
 * pointers and sizes. Host arrays passed in are copied before the call
 * returns; no pointer is retained.
 *
 * Threading: one handle is not thread-safe; calls on a handle are serialized
 * by the caller. Distinct handles are independent. Every call that runs GPU
 * work is synchronous on the handle's own HIP stream.
 *
 * Errors: every int-returning call returns NB_OK (0) or a negative NB_E_*
 * code; nb_last_error() then holds a message. Device-side failures (cell pool
 * exhausted, tree deeper than 64 levels, a bounded wait that ran out) are
 * detected on the device and reported by the nb_step call in which they
 * happened, where the reference only printed them under HOST_DEBUG
 * (GPUBarnesHutNBodySimulation.java:276-290).
 */
#ifndef NBODY_H
#define NBODY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define NB_OK 0
#define NB_E_HIP (-1)   /* HIP runtime error (no device, out of memory, ...) */
#define NB_E_POOL (-2)  /* octree cell pool exhausted (reference: buildtree.cl:112-120) */
#define NB_E_DEPTH (-3) /* tree deeper than NB_MAX_DEPTH (reference: calculateforce.cl:69-73) */
#define NB_E_ARG (-4)   /* invalid argument */
#define NB_E_SYNC (-5)  /* a bounded device-side wait ran out (tree inconsistent) */
#define NB_E_STATE (-6) /* call out of order (e.g. stage before nb_set_state) */
#define NB_E_STACK (-7) /* force traversal stack overflow */
#define NB_E_MODEL (-8) /* learned mode without a model, or a malformed model */

#define NB_MAX_DEPTH 64 /* reference MAXDEPTH, calculateforce.cl:12 */

/* ---- parameters (GPUBarnesHutNBodySimulation.java:104-108, 207-217 and the
 *      compile-time #defines of calculateforce.cl:11-18, integrate.cl:11-12) */
#define NB_FP64_POSITIONS 1u /* leapfrog positions kept in FP64 on the device */
#define NB_FUSED_UPDATE 2u   /* kick/drift fused into the force kernel's epilogue */

#define NB_MODE_PLAIN 0   /* plain Barnes-Hut at `theta` */
#define NB_MODE_LEARNED 1 /* coarse traversal at `coarse_theta` + learned correction */
#define NB_MODE_COARSE 2  /* the coarse traversal alone (no correction): the ablation */

typedef struct nb_params {
    int32_t n_bodies;   /* NBODIES (-D NBODIES, java:212) */
    float theta;        /* opening angle; the reference's THETA is theta^2 (calculateforce.cl:15-16) */
    float softening;    /* the reference's EPSILON is softening^2 (calculateforce.cl:13-14); > 0 */
    float dt;           /* TIMESTEP (calculateforce.cl:18, integrate.cl:12) */
    int32_t vote_width; /* lanes that vote on opening a cell: 16 = the reference's
                           work-group vote (calculateforce.cl:145, WARPSIZE java:47),
                           32, or 64 = one wave64 */
    uint32_t flags;     /* NB_FP64_POSITIONS | NB_FUSED_UPDATE */
    int32_t device;     /* HIP device ordinal */
    float pool_factor;  /* octree cell pool = max(pool_factor*N, 16384) cells */
    int32_t mode;         /* NB_MODE_PLAIN or NB_MODE_LEARNED */
    float coarse_theta;   /* learned mode: opening angle of the coarse traversal's extent
                             test |d| >= b / coarse_theta (b: the cell's extent); in
                             [theta, 1) */
    float check_tol;      /* learned mode: median relative error allowed on the sampled
                             FP64 check before falling back to plain theta; 0 = no check */
    int32_t check_every;  /* learned mode: run the sampled check every k steps (0 = never) */
    int32_t check_samples;/* learned mode: bodies in the sampled check */
    float coarse_alpha;   /* learned mode: relative test m b^3 / s^5 <= coarse_alpha |a_old|
                             of the coarse traversal (0 = off) */
    int32_t check_adapt;  /* learned mode, what a failed check does after its step fell back
                             to plain theta: 0 = plain theta from then on (until a new
                             model); 1 = tighten coarse_alpha for the following steps (aiming
                             at 85 % of check_tol) and keep the learned force, with plain theta
                             from then on after 3 failed checks in a row */
    int32_t reserved[5];
} nb_params;

/* Defaults are the reference's shipped constants: THETA 1.5, EPSILON 0.0025,
 * TIMESTEP 0.025, vote width 16 (calculateforce.cl:11-18). */
void nb_default_params(nb_params *p, int32_t n_bodies);

typedef struct nb_sim nb_sim;

/* new GPUBarnesHutNBodySimulation(mode, nbodies, generator) + init(null):
 * device pick, node-count sizing and buffer allocation
 * (GPUBarnesHutNBodySimulation.java:104-151, 153-181, 219-227). Arguments are
 * validated before any HIP call. */
int nb_create(const nb_params *p, nb_sim **out);

/* Releases the handle and its device memory (no reference counterpart: the
 * reference leaks its OpenCL objects). */
void nb_destroy(nb_sim *sim);

/* UniverseGenerator.generate(0, n, x, y, z, vx, vy, vz, m) followed by the
 * buffer upload of init (UniverseGenerator.java:22, GPUBarnesHutNBodySimulation.java:134-146).
 * Host SoA arrays of n_bodies floats each. Resets the step counter to -1,
 * accelerations to 0 and the running maximum depth to 1 (java:163-170), and
 * the learned mode's control state (nb_control) to a new handle's. */
int nb_set_state(nb_sim *sim, const float *x, const float *y, const float *z,
                 const float *vx, const float *vy, const float *vz, const float *m);

/* AbstractNBodySimulation.step() (AbstractNBodySimulation.java:16), i.e.
 * GPUBarnesHutNBodySimulation.step() (java:249-271): bounding box, tree build,
 * summarize, sort, force, integrate; `nsteps` times, one device error check at
 * the end. */
int nb_step(nb_sim *sim, int32_t nsteps);

/* One stage of step() on its own (1 bbox .. 6 integrate), as the reference's
 * executeSimulationKernel runs one kernel (java:273-291, and
 * TestBoundingBox.java:207 runs the bounding box alone). */
#define NB_STAGE_BBOX 1
#define NB_STAGE_BUILD 2
#define NB_STAGE_SUMMARIZE 3
#define NB_STAGE_SORT 4
#define NB_STAGE_FORCE 5
#define NB_STAGE_INTEGRATE 6
int nb_run_stage(nb_sim *sim, int32_t stage);

/* copyVertices (copyvertices.cl:7-17, launched at java:265-266): writes
 * n_bodies float4 (x, y, z, 1) into pos4 and (vx, vy, vz, 1) into vel4
 * (either may be NULL). Replaces the CL-GL shared VBOs of initGLBuffers
 * (java:229-246): the caller uploads them into its own VBOs. The first call
 * allocates a device staging buffer of 32 n_bodies bytes that later calls
 * reuse (freed by nb_destroy). */
int nb_copy_vertices(nb_sim *sim, float *pos4, float *vel4);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory -- code under HIP graph capture, per-frame render loops, harnesses
 * that forbid hipMalloc: the output of nb_copy_vertices, staged 65,536 bodies
 * at a time through a 2 MiB device array that the library holds on each GPU
 * until the process ends (looked up by nb_create), so the call allocates
 * nothing and has no size limit. Every handle on a GPU shares that array:
 * calls that overlap on one GPU (from different threads) race. */
int nb_copy_vertices_noalloc(nb_sim *sim, float *pos4, float *vel4);

/* getNumberOfBodies() (AbstractNBodySimulation.java:32, java:377-379). */
int32_t nb_num_bodies(const nb_sim *sim);

/* Read-back of the body state by original index (printPosition and the
 * velocity buffer reads of java:305-374). Any pointer may be NULL. */
int nb_get_state(nb_sim *sim, float *x, float *y, float *z, float *vx, float *vy, float *vz);
int nb_get_accel(nb_sim *sim, float *ax, float *ay, float *az);
int nb_get_positions64(nb_sim *sim, double *x, double *y, double *z);

/* Root cell after the bounding box: out4 = (centre x, y, z, radius)
 * (boundingbox.cl:171-189; what TestBoundingBox.java:240-246 reads). */
int nb_get_root(nb_sim *sim, float *out4);

/* sorted[] after the sort stage: body index at each sorted position
 * (sort.cl:57-64). n_bodies int32. */
int nb_get_sorted(nb_sim *sim, int32_t *sorted);

/* Leaf membership after the build stage: per body its octant path from the
 * root (3 bits per level, level 0 in the lowest bits, 21 levels per uint64,
 * 3 words per body) and its depth. Replaces printChildren (java:293-303):
 * cell numbering depends on timing, the octant path does not. */
int nb_get_paths(nb_sim *sim, uint64_t *paths3, int32_t *depth);

/* Per-cell summary after summarize, one entry per cell of the tree (root
 * first, then level by level): body count, centre of mass (FP32) and mass.
 * Arrays of at least `cap` entries (any may be NULL); *ncells receives the
 * number of cells. */
int nb_get_cells(nb_sim *sim, int32_t cap, int32_t *ncells, int32_t *count, float *x,
                 float *y, float *z, float *m);

typedef struct nb_counters {
    int64_t step;          /* the reference's _step counter */
    int64_t cells;         /* cells of the last build */
    int64_t max_depth;     /* running maximum depth (reference _maxDepth) */
    int64_t interactions;  /* body-node interactions of the last force stage */
    int64_t error;         /* device error word */
    int64_t fallbacks;     /* learned mode: steps that fell back to plain theta */
    double stage_ms[8];    /* per-step device time of each stage over the last nb_step call (timing on) */
    double last_check_median; /* learned mode: last sampled median relative error */
    int64_t live_cells;    /* cells linked into the tree by the last build */
    int64_t cell_interactions; /* learned and coarse modes: the part of `interactions` that
                                  were cells (monopole + quadrupole) */
    double check_ms;       /* learned mode: device time of the sampled FP64 checks (direct sum
                              and gather) since nb_create, cumulative */
    int64_t checks;        /* learned mode: sampled checks run since nb_create */
    double coarse_alpha;   /* learned mode: the relative test's alpha in effect (check_adapt) */
    int64_t large_buckets; /* builds from the last tree's order: buckets too large to sort on
                              chip (sorted through device memory), cumulative since the last
                              nb_set_state */
} nb_counters;
int nb_get_counters(nb_sim *sim, nb_counters *out);

/* ---- force accuracy (no reference counterpart: the reference has no error
 *      measurement; these back the sampled FP64 check of the learned mode and
 *      the error figures of the benchmark) ---------------------------------- */

/* Accelerations that the configured force method gives for the current
 * positions (bounding box .. force, without advancing the bodies). */
int nb_probe_accel(nb_sim *sim, float *ax, float *ay, float *az);

/* FP64 direct-sum accelerations (G = 1, same softening) of the bodies idx[0..ns)
 * at the current positions, computed on the device; out3 holds 3*ns doubles. */
int nb_direct_accel(nb_sim *sim, const int32_t *idx, int32_t ns, double *out3);

/* Relative force error of the configured method against the FP64 direct sum
 * on nsamples distinct bodies drawn with `seed`: stats4 = (median, 90th
 * percentile, maximum, mean) of |a - a_direct| / |a_direct|. */
int nb_force_error(nb_sim *sim, int32_t nsamples, uint64_t seed, double *stats4);

/* ---- learned coarse-theta correction (no reference counterpart) ----------
 *
 * The coarse traversal accepts a cell of extent b (the largest distance of its
 * bodies from its centre of mass) and mass m at softened distance s when, for
 * all 64 bodies of the vote group, |d| >= b / coarse_theta and
 * m b^3 / s^5 <= coarse_alpha |a_old| (a_old: the body's acceleration of the
 * previous step; this test is off until there is one), and adds, to the
 * monopole of every accepted cell, its quadrupole under the softened kernel.
 * It also emits 24 features per body in a frame built from the body's own
 * coarse acceleration a (e1 = a/|a|, e2 from the quadrupole part, e3 = e1 x e2).
 * A per-body MLP 24 -> 64 -> 64 -> 3 (tanh-GELU) maps the normalized features
 * to the residual (a_fine - a) in that frame over |a|, where a_fine is the
 * same traversal with the force stage's opening table at `theta`. Inference runs with BF16 weights and activations
 * on the matrix cores with FP32 accumulation; training runs in FP32 with
 * Adam on pairs sampled from the simulation itself.
 *
 * Model blob (NB_MODEL_FLOATS floats): the NB_MLP_PARAMS network parameters
 * W1[24][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3] (W[i][j] maps input
 * i to output j), then the feature means mu_x[24] and standard deviations
 * sd_x[24], then the target means mu_y[3] and standard deviations sd_y[3]:
 * residual = MLP(clip((x - mu_x) / sd_x)) * sd_y + mu_y, where clip limits each
 * normalized feature to [-8, 8] (training clips its normalized targets the same
 * way). */
#define NB_MLP_IN 24
#define NB_MLP_HIDDEN 64
#define NB_MLP_OUT 3
#define NB_MLP_PARAMS 5955
#define NB_MODEL_FLOATS (NB_MLP_PARAMS + 2 * NB_MLP_IN + 2 * NB_MLP_OUT)

typedef struct nb_train_params {
    int32_t steps;          /* simulation steps that generate pairs (the bodies advance with plain theta) */
    int32_t groups_per_step;/* vote groups of 64 bodies sampled per step */
    int32_t iters;          /* Adam iterations */
    int32_t batch;          /* minibatch size */
    float lr;               /* Adam learning rate (cosine decay to lr_final) */
    float lr_final;
    float beta1, beta2, adam_eps;
    uint64_t seed;
    int32_t reserved[6];
} nb_train_params;

typedef struct nb_train_report {
    int32_t samples;        /* training pairs collected */
    float loss_first;       /* minibatch loss of the first and of the last iteration (normalized units) */
    float loss_last;
    double gen_ms;          /* pair generation, including the plain steps */
    double train_ms;        /* the Adam iterations */
    int64_t reserved[4];
} nb_train_report;

void nb_default_train_params(nb_train_params *tp);

/* Generates pairs over tp->steps steps, trains the network from a seeded
 * initialization and installs it; the simulation has advanced tp->steps
 * steps with plain theta. rep may be NULL. */
int nb_train(nb_sim *sim, const nb_train_params *tp, nb_train_report *rep);

/* Installs / reads the model blob (NB_MODEL_FLOATS floats). */
int nb_set_model(nb_sim *sim, const float *model, int32_t count);
int nb_get_model(nb_sim *sim, float *model, int32_t count);

/* Coarse-traversal features at the current positions (bounding box .. sort
 * are run first): feat (24 per body), acc3 (the coarse acceleration, 3 per
 * body) and e2 (3 per body), all by original body index; any may be NULL. */
int nb_coarse_features(nb_sim *sim, float *feat, float *acc3, float *e2);

/* The network alone, on host data. nb_mlp_forward: raw features x[n][24] ->
 * residual y[n][3] through the installed model on the matrix-core path.
 * nb_mlp_loss_grad: for the current FP32 parameters, normalized inputs
 * x[n][24] and targets t[n][3], the mean squared error over n*3 outputs and
 * its gradient (NB_MLP_PARAMS floats) from the FP32 training kernels.
 * nb_mlp_adam: `iters` full-batch Adam steps on (x, t) with lr and the
 * default betas, starting from the current parameters and zeroed moments;
 * params receives the result. */
int nb_mlp_forward(nb_sim *sim, const float *x, int32_t n, float *y);
int nb_mlp_loss_grad(nb_sim *sim, const float *x, const float *t, int32_t n, float *loss,
                     float *grad);
int nb_mlp_adam(nb_sim *sim, const float *x, const float *t, int32_t n, int32_t iters, float lr,
                float *params);

/* ---- diagnostics and checkpoints (SURVEY-listed follow-ups; the reference's
 *      printEnergy, GPUBarnesHutNBodySimulation.java:305-365, counts every pair
 *      twice and ignores the softening) -------------------------------------- */

/* FP64 sums at the current state (positions in FP64 when NB_FP64_POSITIONS):
 * out12[0] kinetic energy, out12[1] potential energy
 * -1/2 sum_i sum_{j != i} m_i m_j / sqrt(|x_i - x_j|^2 + eps^2) over the FP32
 * positions the force sees -- exact (all N bodies, O(N^2) on the device) when
 * pe_samples <= 0 or >= N, otherwise estimated from pe_samples distinct bodies
 * drawn with `seed` and scaled by N / pe_samples; out12[2..4] momentum,
 * out12[5..7] angular momentum about the origin, out12[8..10] centre of mass,
 * out12[11] total mass. */
int nb_diagnostics(nb_sim *sim, int32_t pe_samples, uint64_t seed, double *out12);

/* Masses by original body index (the missing read-back for a checkpoint). */
int nb_get_masses(nb_sim *sim, float *m);

/* Resume from a checkpoint: positions (FP64; rounded to FP32 for the tree),
 * velocities, last accelerations and masses by original body index, and the
 * step counter and running maximum depth (nb_get_counters). In plain and
 * coarse modes the next nb_step continues exactly as the interrupted run
 * would have; the learned mode also needs its model and nb_control back
 * (nb_set_model, then nb_set_control). */
int nb_restore_state(nb_sim *sim, const double *x, const double *y, const double *z, const float *vx,
                     const float *vy, const float *vz, const float *ax, const float *ay, const float *az,
                     const float *m, int32_t step, int32_t max_depth);

/* Learned-mode control state: what, beside the bodies (nb_restore_state) and
 * the model (nb_get_model / nb_set_model), decides how the next steps of a
 * learned run go -- the sampled check's schedule and sample generator, the
 * relative test's alpha after check_adapt tightened it, and a fallback to
 * plain theta. A checkpoint saves it after the model; a resume calls
 * nb_restore_state, nb_set_model, then nb_set_control (nb_set_state and
 * nb_set_model put it back to a new handle's values). No reference
 * counterpart: the reference has no learned mode and no checkpoint. */
typedef struct nb_control {
    int64_t steps_since_model; /* learned steps since the model was installed: a step
                                  runs the check when this is a multiple of check_every */
    uint64_t check_seed;       /* state of the generator that draws the checked bodies */
    float coarse_alpha;        /* the relative test's alpha in effect */
    int32_t fell_back;         /* 1: a failed check switched the force to plain theta */
    int32_t fail_streak;       /* failed checks in a row */
    int32_t reserved0;
    int64_t fallbacks;         /* counters carried into nb_get_counters: fallback steps, */
    int64_t checks;            /* checks run, */
    double last_check_median;  /* and the last check's median relative error */
    int64_t reserved[4];
} nb_control;
int nb_get_control(nb_sim *sim, nb_control *out);
int nb_set_control(nb_sim *sim, const nb_control *in);

/* Per-stage HIP-event timing of the following nb_step calls (0 = off). */
int nb_set_timing(nb_sim *sim, int32_t enabled);

/* Message for the last failing call on this handle ("" if none); with a NULL
 * handle, the message of the last nb_create that failed on this thread. */
const char *nb_last_error(const nb_sim *sim);

/* Static text for a status code. */
const char *nb_status_string(int code);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_H */
