/*
 * nbody_jni_core.c -- the JNI glue's work in plain C (see nbody_jni_core.h).
 * Checks what libnbody.so cannot see from a raw pointer -- the length of a
 * Java array, the capacity of a direct buffer -- before handing the pointers
 * on, so a short array is an IllegalArgumentException instead of a read past
 * its end.
 */
#include "nbody_jni_core.h"

#include <stddef.h>
#include <string.h>

static int bad(const char **why, const char *msg)
{
    if (why) *why = msg;
    return NB_E_ARG;
}

static void clear(const char **why)
{
    if (why) *why = NULL;
}

int nbj_params(int32_t n, const float *f, int32_t nf, const int32_t *i, int32_t ni, nb_params *out,
               const char **why)
{
    clear(why);
    if (!out) return bad(why, "no parameter record");
    if (n < 1) return bad(why, "nbodies must be positive");
    if (!f || nf != NBJ_NF) return bad(why, "float parameter array must hold 7 values");
    if (!i || ni != NBJ_NI) return bad(why, "int parameter array must hold 7 values");
    nb_default_params(out, n);
    out->theta = f[0];
    out->softening = f[1];
    out->dt = f[2];
    out->pool_factor = f[3];
    out->coarse_theta = f[4];
    out->check_tol = f[5];
    out->coarse_alpha = f[6];
    out->vote_width = i[0];
    out->flags = (uint32_t)i[1];
    out->device = i[2];
    out->mode = i[3];
    out->check_every = i[4];
    out->check_samples = i[5];
    out->check_adapt = i[6];
    return NB_OK;
}

int nbj_create(int32_t n, const float *f, int32_t nf, const int32_t *i, int32_t ni, nb_sim **out,
               const char **why)
{
    nb_params p;
    int r = nbj_params(n, f, nf, i, ni, &p, why);
    if (r) return r;
    if (!out) return bad(why, "no handle slot");
    return nb_create(&p, out);
}

static int check_arrays(nb_sim *sim, float *const *arr, const int32_t *len, int k, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!arr || !len) return bad(why, "no state arrays");
    const int32_t n = nb_num_bodies(sim);
    for (int j = 0; j < k; ++j) {
        if (!arr[j]) return bad(why, "state array is null");
        if (len[j] < n) return bad(why, "state array shorter than the number of bodies");
    }
    return NB_OK;
}

int nbj_set_state(nb_sim *sim, float *const arr[7], const int32_t len[7], const char **why)
{
    int r = check_arrays(sim, arr, len, 7, why);
    if (r) return r;
    return nb_set_state(sim, arr[0], arr[1], arr[2], arr[3], arr[4], arr[5], arr[6]);
}

int nbj_get_state(nb_sim *sim, float *const arr[6], const int32_t len[6], const char **why)
{
    int r = check_arrays(sim, arr, len, 6, why);
    if (r) return r;
    return nb_get_state(sim, arr[0], arr[1], arr[2], arr[3], arr[4], arr[5]);
}

int nbj_step(nb_sim *sim, int32_t steps, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (steps < 0) return bad(why, "negative step count");
    return nb_step(sim, steps);
}

int nbj_copy_vertices(nb_sim *sim, float *pos4, int64_t pos_floats, float *vel4, int64_t vel_floats,
                      const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    const int64_t need = 4 * (int64_t)nb_num_bodies(sim);
    if (pos4 && pos_floats < need) return bad(why, "position buffer holds fewer than 4 floats per body");
    if (vel4 && vel_floats < need) return bad(why, "velocity buffer holds fewer than 4 floats per body");
    return nb_copy_vertices(sim, pos4, vel4);
}

int nbj_train(nb_sim *sim, const float *tf, int32_t ntf, const int32_t *ti, int32_t nti, int64_t seed,
              double *rep, int32_t nrep, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!tf || ntf != NBJ_TF) return bad(why, "float training array must hold 5 values");
    if (!ti || nti != NBJ_TI) return bad(why, "int training array must hold 4 values");
    if (rep && nrep < NBJ_TREP) return bad(why, "report array must hold 5 values");
    nb_train_params tp;
    nb_default_train_params(&tp);
    tp.lr = tf[0];
    tp.lr_final = tf[1];
    tp.beta1 = tf[2];
    tp.beta2 = tf[3];
    tp.adam_eps = tf[4];
    tp.steps = ti[0];
    tp.groups_per_step = ti[1];
    tp.iters = ti[2];
    tp.batch = ti[3];
    tp.seed = (uint64_t)seed;
    nb_train_report r;
    int code = nb_train(sim, &tp, &r);
    if (code == NB_OK && rep) {
        rep[0] = r.samples;
        rep[1] = r.loss_first;
        rep[2] = r.loss_last;
        rep[3] = r.gen_ms;
        rep[4] = r.train_ms;
    }
    return code;
}

int nbj_set_model(nb_sim *sim, const float *model, int32_t len, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!model || len != NB_MODEL_FLOATS) return bad(why, "model array must hold NB_MODEL_FLOATS floats");
    return nb_set_model(sim, model, len);
}

int nbj_get_model(nb_sim *sim, float *model, int32_t len, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!model || len < NB_MODEL_FLOATS) return bad(why, "model array must hold NB_MODEL_FLOATS floats");
    return nb_get_model(sim, model, len);
}

int nbj_counters(nb_sim *sim, int64_t *l, int32_t nl, double *d, int32_t nd, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!l || nl < NBJ_CL || !d || nd < NBJ_CD) return bad(why, "counter arrays too short");
    nb_counters c;
    int r = nb_get_counters(sim, &c);
    if (r) return r;
    const int64_t lv[NBJ_CL] = {c.step,  c.cells,      c.max_depth,         c.interactions, c.error,
                                c.fallbacks, c.live_cells, c.cell_interactions, c.checks};
    memcpy(l, lv, sizeof lv);
    for (int k = 0; k < 6; ++k) d[k] = c.stage_ms[k];
    d[6] = c.last_check_median;
    d[7] = c.check_ms;
    d[8] = c.coarse_alpha;
    return NB_OK;
}

int nbj_get_control(nb_sim *sim, int64_t *l, int32_t nl, double *d, int32_t nd, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!l || nl < NBJ_KL || !d || nd < NBJ_KD) return bad(why, "control arrays too short");
    nb_control k;
    int r = nb_get_control(sim, &k);
    if (r) return r;
    l[0] = k.steps_since_model;
    l[1] = (int64_t)k.check_seed;
    l[2] = k.fell_back;
    l[3] = k.fail_streak;
    l[4] = k.fallbacks;
    l[5] = k.checks;
    d[0] = k.coarse_alpha;
    d[1] = k.last_check_median;
    return NB_OK;
}

int nbj_set_control(nb_sim *sim, const int64_t *l, int32_t nl, const double *d, int32_t nd, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!l || nl != NBJ_KL || !d || nd != NBJ_KD) return bad(why, "control arrays must hold 6 longs and 2 doubles");
    if (l[2] < 0 || l[2] > 1 || l[3] < 0 || l[3] > INT32_MAX) return bad(why, "invalid control state");
    nb_control k;
    memset(&k, 0, sizeof k);
    k.steps_since_model = l[0];
    k.check_seed = (uint64_t)l[1];
    k.fell_back = (int32_t)l[2];
    k.fail_streak = (int32_t)l[3];
    k.fallbacks = l[4];
    k.checks = l[5];
    k.coarse_alpha = (float)d[0];
    k.last_check_median = d[1];
    return nb_set_control(sim, &k);
}

int nbj_force_error(nb_sim *sim, int32_t nsamples, int64_t seed, double *stats, int32_t nstats, const char **why)
{
    clear(why);
    if (!sim) return bad(why, "simulation not initialized");
    if (!stats || nstats < 4) return bad(why, "statistics array must hold 4 values");
    return nb_force_error(sim, nsamples, (uint64_t)seed, stats);
}

void nbj_destroy(nb_sim *sim) { nb_destroy(sim); }

const char *nbj_message(const nb_sim *sim, int code, const char *why)
{
    if (why && why[0]) return why;
    const char *msg = nb_last_error(sim);
    return (msg && msg[0]) ? msg : nb_status_string(code);
}
