/*
 * nbody_jni_core.h -- what the JNI glue of HipBarnesHutNBodySimulation does,
 * in plain C: argument checks on the Java arrays and buffers (lengths,
 * capacities, NULL), the packing of the Java-side parameter arrays into
 * nb_params / nb_train_params, and the calls into include/nbody.h.
 * nbody_jni.c only moves data between the JVM and these functions, so the
 * whole call sequence of the Java class can be run and tested from C without
 * a JDK.
 *
 * Every function returns NB_OK or an NB_E_* code. A check done here (not in
 * libnbody.so) stores a static message in *why; otherwise *why is NULL and
 * nb_last_error(sim) holds the library's message (nbj_message picks one).
 */
#ifndef NBODY_JNI_CORE_H
#define NBODY_JNI_CORE_H

#include <stdint.h>

#include "nbody.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Float parameters, in this order (HipBarnesHutNBodySimulation.Params.floats()):
 * theta, softening, dt, pool_factor, coarse_theta, check_tol, coarse_alpha */
#define NBJ_NF 7
/* Int parameters: vote_width, flags, device, mode, check_every, check_samples, check_adapt */
#define NBJ_NI 7
/* Training, floats: lr, lr_final, beta1, beta2, adam_eps; ints: steps,
 * groups_per_step, iters, batch (the seed is a separate long) */
#define NBJ_TF 5
#define NBJ_TI 4
/* Training report: samples, loss_first, loss_last, gen_ms, train_ms */
#define NBJ_TREP 5
/* Counters, longs: step, cells, max_depth, interactions, error, fallbacks,
 * live_cells, cell_interactions, checks; doubles: stage_ms[0..5],
 * last_check_median, check_ms, coarse_alpha */
#define NBJ_CL 9
#define NBJ_CD 9
/* Control state, longs: steps_since_model, check_seed, fell_back, fail_streak,
 * fallbacks, checks; doubles: coarse_alpha, last_check_median */
#define NBJ_KL 6
#define NBJ_KD 2

/* nb_default_params(n), then the Java arrays on top of it. */
int nbj_params(int32_t n, const float *f, int32_t nf, const int32_t *i, int32_t ni, nb_params *out,
               const char **why);
int nbj_create(int32_t n, const float *f, int32_t nf, const int32_t *i, int32_t ni, nb_sim **out,
               const char **why);
/* arr: x, y, z, vx, vy, vz, m with their Java lengths; each must hold n floats. */
int nbj_set_state(nb_sim *sim, float *const arr[7], const int32_t len[7], const char **why);
int nbj_get_state(nb_sim *sim, float *const arr[6], const int32_t len[6], const char **why);
int nbj_step(nb_sim *sim, int32_t steps, const char **why);
/* pos4 / vel4: direct buffers (either NULL), capacities in floats, >= 4 n each. */
int nbj_copy_vertices(nb_sim *sim, float *pos4, int64_t pos_floats, float *vel4, int64_t vel_floats,
                      const char **why);
int nbj_train(nb_sim *sim, const float *tf, int32_t ntf, const int32_t *ti, int32_t nti, int64_t seed,
              double *rep, int32_t nrep, const char **why);
int nbj_set_model(nb_sim *sim, const float *model, int32_t len, const char **why);
int nbj_get_model(nb_sim *sim, float *model, int32_t len, const char **why);
int nbj_counters(nb_sim *sim, int64_t *l, int32_t nl, double *d, int32_t nd, const char **why);
int nbj_get_control(nb_sim *sim, int64_t *l, int32_t nl, double *d, int32_t nd, const char **why);
int nbj_set_control(nb_sim *sim, const int64_t *l, int32_t nl, const double *d, int32_t nd, const char **why);
/* stats4 = (median, 90th percentile, maximum, mean) relative force error. */
int nbj_force_error(nb_sim *sim, int32_t nsamples, int64_t seed, double *stats, int32_t nstats,
                    const char **why);
void nbj_destroy(nb_sim *sim);

/* The text an exception carries for a failed call. */
const char *nbj_message(const nb_sim *sim, int code, const char *why);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_JNI_CORE_H */
