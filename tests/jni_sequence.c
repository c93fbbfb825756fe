/*
 * jni_sequence.c -- the call sequence of HipBarnesHutNBodySimulation.java,
 * replayed through nbody_jni_core.c (what nbody_jni.c calls) against
 * libnbody.so, without a JVM.
 *
 *   jni_sequence args              argument checks only (no GPU call)
 *   jni_sequence gpu UNIVERSE      the full sequence on device 0; UNIVERSE is a
 *                                  file: int32 n, then x, y, z, vx, vy, vz, m
 *                                  (n float32 each, little endian)
 *
 * The sequence follows the Java class: init (nativeCreate with Params'
 * arrays, then nativeSetState with the generated universe), initGLBuffers
 * (two direct buffers of 4 n floats), step() in GL_INTEROP mode
 * (nativeStep, nativeCopyVertices), counters, dispose (nativeDestroy); then
 * the learned mode: train, getModel / setModel onto a second handle, steps,
 * control state, forceError. Prints one line per check and
 * "JNI_SEQUENCE_OK" at the end; exits 1 at the first failed check.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nbody_jni_core.h"

static int g_checks = 0;

#define CHECK(cond, ...)                                                                                  \
    do {                                                                                                  \
        ++g_checks;                                                                                       \
        if (!(cond)) {                                                                                    \
            fprintf(stderr, "FAILED %s:%d: %s -- ", __FILE__, __LINE__, #cond);                           \
            fprintf(stderr, __VA_ARGS__);                                                                 \
            fprintf(stderr, "\n");                                                                        \
            exit(1);                                                                                      \
        }                                                                                                 \
    } while (0)

/* HipBarnesHutNBodySimulation.Params defaults, as floats() / ints() pack them */
static void java_defaults(float f[NBJ_NF], int32_t i[NBJ_NI])
{
    const float d_f[NBJ_NF] = {sqrtf(1.5f), 0.05f, 0.025f, 2.0f, 0.95f, 1e-3f, 0.02f};
    const int32_t d_i[NBJ_NI] = {16, 0, 0, 0, 0, 4096, 0};
    memcpy(f, d_f, sizeof d_f);
    memcpy(i, d_i, sizeof d_i);
}

static int args_only(void)
{
    float f[NBJ_NF];
    int32_t i[NBJ_NI];
    java_defaults(f, i);
    const char *why = NULL;
    nb_params p, d;
    /* the Java defaults are the library's defaults */
    CHECK(nbj_params(1000, f, NBJ_NF, i, NBJ_NI, &p, &why) == NB_OK && why == NULL, "params");
    nb_default_params(&d, 1000);
    CHECK(memcmp(&p, &d, sizeof p) == 0, "Java Params defaults differ from nb_default_params");
    /* packing order */
    const float f2[NBJ_NF] = {0.5f, 0.01f, 1e-3f, 3.0f, 0.9f, 2e-3f, 0.015f};
    const int32_t i2[NBJ_NI] = {64, 3, 0, 1, 10, 1024, 1};
    CHECK(nbj_params(4096, f2, NBJ_NF, i2, NBJ_NI, &p, &why) == NB_OK, "params 2");
    CHECK(p.n_bodies == 4096 && p.theta == 0.5f && p.softening == 0.01f && p.dt == 1e-3f && p.pool_factor == 3.0f &&
              p.coarse_theta == 0.9f && p.check_tol == 2e-3f && p.coarse_alpha == 0.015f,
          "float packing");
    CHECK(p.vote_width == 64 && p.flags == 3u && p.device == 0 && p.mode == 1 && p.check_every == 10 &&
              p.check_samples == 1024 && p.check_adapt == 1,
          "int packing");
    /* wrong array sizes and null handles are rejected before any HIP call */
    nb_sim *sim = NULL;
    CHECK(nbj_create(1000, f, 6, i, NBJ_NI, &sim, &why) == NB_E_ARG && why && sim == NULL, "short float params");
    CHECK(nbj_create(1000, f, NBJ_NF, i, 8, &sim, &why) == NB_E_ARG && why, "long int params");
    CHECK(nbj_create(1000, NULL, NBJ_NF, i, NBJ_NI, &sim, &why) == NB_E_ARG && why, "null params");
    CHECK(nbj_create(0, f, NBJ_NF, i, NBJ_NI, &sim, &why) == NB_E_ARG && why, "no bodies");
    float *arr[7] = {0};
    int32_t len[7] = {0};
    CHECK(nbj_set_state(NULL, arr, len, &why) == NB_E_ARG && why, "null handle set_state");
    CHECK(nbj_get_state(NULL, arr, len, &why) == NB_E_ARG && why, "null handle get_state");
    CHECK(nbj_step(NULL, 1, &why) == NB_E_ARG && why, "null handle step");
    CHECK(nbj_copy_vertices(NULL, NULL, 0, NULL, 0, &why) == NB_E_ARG && why, "null handle copy");
    CHECK(nbj_train(NULL, NULL, 0, NULL, 0, 1, NULL, 0, &why) == NB_E_ARG && why, "null handle train");
    CHECK(nbj_set_model(NULL, NULL, 0, &why) == NB_E_ARG && why, "null handle set_model");
    CHECK(nbj_get_model(NULL, NULL, 0, &why) == NB_E_ARG && why, "null handle get_model");
    int64_t l[NBJ_CL];
    double dd[NBJ_CD];
    CHECK(nbj_counters(NULL, l, NBJ_CL, dd, NBJ_CD, &why) == NB_E_ARG && why, "null handle counters");
    CHECK(nbj_get_control(NULL, l, NBJ_KL, dd, NBJ_KD, &why) == NB_E_ARG && why, "null handle get_control");
    CHECK(nbj_set_control(NULL, l, NBJ_KL, dd, NBJ_KD, &why) == NB_E_ARG && why, "null handle set_control");
    CHECK(nbj_force_error(NULL, 16, 1, dd, 4, &why) == NB_E_ARG && why, "null handle force_error");
    CHECK(strcmp(nbj_message(NULL, NB_E_ARG, "local"), "local") == 0, "message");
    nbj_destroy(NULL);
    printf("args: %d checks\n", g_checks);
    return 0;
}

static float *read_universe(const char *path, int32_t *n_out)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return NULL;
    int32_t n = 0;
    if (fread(&n, 4, 1, fp) != 1 || n < 1) {
        fclose(fp);
        return NULL;
    }
    float *buf = (float *)malloc((size_t)7 * n * sizeof(float));
    if (buf && fread(buf, sizeof(float), (size_t)7 * n, fp) != (size_t)7 * n) {
        free(buf);
        buf = NULL;
    }
    fclose(fp);
    *n_out = n;
    return buf;
}

static void vertices_match(nb_sim *sim, int32_t n, const float *pos4, const float *vel4)
{
    float *st = (float *)malloc((size_t)6 * n * sizeof(float));
    float *a[6];
    int32_t len[6];
    for (int j = 0; j < 6; ++j) {
        a[j] = st + (size_t)j * n;
        len[j] = n;
    }
    const char *why = NULL;
    CHECK(nbj_get_state(sim, a, len, &why) == NB_OK, "get_state: %s", nbj_message(sim, -1, why));
    for (int32_t b = 0; b < n; ++b) {
        const float *p = pos4 + 4 * (size_t)b, *v = vel4 + 4 * (size_t)b;
        if (memcmp(&p[0], &a[0][b], 4) || memcmp(&p[1], &a[1][b], 4) || memcmp(&p[2], &a[2][b], 4) || p[3] != 1.0f ||
            memcmp(&v[0], &a[3][b], 4) || memcmp(&v[1], &a[4][b], 4) || memcmp(&v[2], &a[5][b], 4) || v[3] != 1.0f) {
            CHECK(0, "vertex %d differs from the state", b);
        }
    }
    free(st);
}

static int full_sequence(const char *path)
{
    int32_t n = 0;
    float *u = read_universe(path, &n);
    CHECK(u != NULL, "cannot read %s", path);
    float *arr[7];
    int32_t len[7];
    for (int j = 0; j < 7; ++j) {
        arr[j] = u + (size_t)j * n;
        len[j] = n;
    }
    const char *why = NULL;

    /* ---- new HipBarnesHutNBodySimulation(GL_INTEROP, n, generator); init(gl) */
    float f[NBJ_NF];
    int32_t i[NBJ_NI];
    java_defaults(f, i);
    nb_sim *sim = NULL;
    int r = nbj_create(n, f, NBJ_NF, i, NBJ_NI, &sim, &why);
    CHECK(r == NB_OK && sim, "create: %s", nbj_message(NULL, r, why));
    r = nbj_set_state(sim, arr, len, &why);
    CHECK(r == NB_OK, "set_state: %s", nbj_message(sim, r, why));
    printf("init: %d bodies, reference constants\n", n);

    /* ---- initGLBuffers: two direct FloatBuffers of 4 n floats */
    float *pos4 = (float *)malloc((size_t)16 * n), *vel4 = (float *)malloc((size_t)16 * n);
    /* ---- step() in GL_INTEROP mode, three frames */
    for (int k = 0; k < 3; ++k) {
        memset(pos4, 0xff, (size_t)16 * n);
        memset(vel4, 0xff, (size_t)16 * n);
        r = nbj_step(sim, 1, &why);
        CHECK(r == NB_OK, "step: %s", nbj_message(sim, r, why));
        r = nbj_copy_vertices(sim, pos4, 4 * (int64_t)n, vel4, 4 * (int64_t)n, &why);
        CHECK(r == NB_OK, "copy_vertices: %s", nbj_message(sim, r, why));
        vertices_match(sim, n, pos4, vel4);
    }
    printf("step: 3 frames, vertices = state with w = 1\n");

    /* ---- what the glue rejects on a live handle */
    len[4] = n - 1;
    r = nbj_set_state(sim, arr, len, &why);
    CHECK(r == NB_E_ARG && why, "short array accepted");
    len[4] = n;
    arr[6] = NULL;
    CHECK(nbj_set_state(sim, arr, len, &why) == NB_E_ARG && why, "null array accepted");
    arr[6] = u + (size_t)6 * n;
    CHECK(nbj_copy_vertices(sim, pos4, 4 * (int64_t)n - 1, vel4, 4 * (int64_t)n, &why) == NB_E_ARG && why,
          "short vertex buffer accepted");
    CHECK(nbj_copy_vertices(sim, NULL, 0, vel4, 4 * (int64_t)n, &why) == NB_OK, "positions NULL");
    CHECK(nbj_step(sim, -1, &why) == NB_E_ARG && why, "negative steps accepted");
    printf("glue rejects short arrays, short buffers, negative steps\n");

    int64_t l[NBJ_CL];
    double d[NBJ_CD];
    r = nbj_counters(sim, l, NBJ_CL, d, NBJ_CD, &why);
    CHECK(r == NB_OK && l[0] == 2 && l[3] > 0 && l[4] == 0, "counters: step %lld interactions %lld error %lld",
          (long long)l[0], (long long)l[3], (long long)l[4]);
    /* ---- dispose() */
    nbj_destroy(sim);
    printf("counters: step %lld, %lld interactions; dispose\n", (long long)l[0], (long long)l[3]);

    /* ---- the learned mode through the four-argument constructor */
    const float fl[NBJ_NF] = {0.4f, 0.01f, 1e-3f, 2.0f, 0.9f, 10.0f, 0.02f};
    const int32_t il[NBJ_NI] = {64, 2 /* FUSED_UPDATE */, 0, 1 /* MODE_LEARNED */, 2, 256, 1};
    nb_sim *a = NULL, *b = NULL;
    r = nbj_create(n, fl, NBJ_NF, il, NBJ_NI, &a, &why);
    CHECK(r == NB_OK, "create learned: %s", nbj_message(NULL, r, why));
    r = nbj_create(n, fl, NBJ_NF, il, NBJ_NI, &b, &why);
    CHECK(r == NB_OK, "create learned: %s", nbj_message(NULL, r, why));
    CHECK(nbj_set_state(a, arr, len, &why) == NB_OK, "set_state a");
    CHECK(nbj_step(a, 1, &why) == NB_E_MODEL, "learned step without a model accepted");
    const float tf[NBJ_TF] = {2e-3f, 1e-4f, 0.9f, 0.999f, 1e-8f};
    const int32_t ti[NBJ_TI] = {3, 64, 50, 1024};
    double rep[NBJ_TREP];
    r = nbj_train(a, tf, NBJ_TF, ti, NBJ_TI, 7, rep, NBJ_TREP, &why);
    CHECK(r == NB_OK, "train: %s", nbj_message(a, r, why));
    CHECK(rep[0] == 3 * 64 * 64 && isfinite(rep[1]) && isfinite(rep[2]), "train report");
    printf("train: %g pairs, loss %.4f -> %.4f\n", rep[0], rep[1], rep[2]);
    float *model = (float *)malloc(sizeof(float) * NB_MODEL_FLOATS);
    CHECK(nbj_get_model(a, model, NB_MODEL_FLOATS, &why) == NB_OK, "get_model");
    CHECK(nbj_set_model(b, model, NB_MODEL_FLOATS - 1, &why) == NB_E_ARG && why, "short model accepted");
    /* the same universe and model on both handles: identical runs */
    CHECK(nbj_set_state(a, arr, len, &why) == NB_OK, "set_state a");
    CHECK(nbj_set_state(b, arr, len, &why) == NB_OK, "set_state b");
    CHECK(nbj_set_model(b, model, NB_MODEL_FLOATS, &why) == NB_OK, "set_model b");
    for (int k = 0; k < 4; ++k) {
        CHECK(nbj_step(a, 1, &why) == NB_OK, "step a: %s", nbj_message(a, -1, why));
    }
    CHECK(nbj_step(b, 4, &why) == NB_OK, "steps b: %s", nbj_message(b, -1, why));
    {
        float *sa = (float *)malloc((size_t)12 * n * sizeof(float));
        float *pa[6], *pb[6];
        int32_t ln[6];
        for (int j = 0; j < 6; ++j) {
            pa[j] = sa + (size_t)j * n;
            pb[j] = sa + (size_t)(6 + j) * n;
            ln[j] = n;
        }
        CHECK(nbj_get_state(a, pa, ln, &why) == NB_OK && nbj_get_state(b, pb, ln, &why) == NB_OK, "get_state");
        CHECK(memcmp(sa, sa + (size_t)6 * n, (size_t)6 * n * sizeof(float)) == 0,
              "one call of 4 steps differs from 4 single steps");
        free(sa);
    }
    int64_t ka[NBJ_KL], kb[NBJ_KL];
    double da[NBJ_KD], db[NBJ_KD];
    CHECK(nbj_get_control(a, ka, NBJ_KL, da, NBJ_KD, &why) == NB_OK, "get_control");
    CHECK(ka[0] == 4 && ka[5] == 2 && ka[2] == 0, "control: steps %lld checks %lld fell_back %lld", (long long)ka[0],
          (long long)ka[5], (long long)ka[2]);
    CHECK(nbj_set_control(b, ka, NBJ_KL, da, NBJ_KD, &why) == NB_OK, "set_control");
    CHECK(nbj_get_control(b, kb, NBJ_KL, db, NBJ_KD, &why) == NB_OK, "get_control b");
    CHECK(memcmp(ka, kb, sizeof ka) == 0 && (float)da[0] == (float)db[0], "control round trip");
    double stats[4];
    r = nbj_force_error(a, 512, 3, stats, 4, &why);
    CHECK(r == NB_OK && stats[0] > 0.0 && stats[0] < 1e-1 && stats[1] >= stats[0], "force_error: %s",
          nbj_message(a, r, why));
    printf("learned: 4 steps identical on two handles, control round trip, median error %.3e\n", stats[0]);
    nbj_destroy(a);
    nbj_destroy(b);
    free(model);
    free(pos4);
    free(vel4);
    free(u);
    return 0;
}

int main(int argc, char **argv)
{
    int r = 1;
    if (argc >= 2 && strcmp(argv[1], "args") == 0)
        r = args_only();
    else if (argc >= 3 && strcmp(argv[1], "gpu") == 0)
        r = args_only() || full_sequence(argv[2]);
    else {
        fprintf(stderr, "usage: %s args | gpu UNIVERSE\n", argv[0]);
        return 2;
    }
    if (r == 0) printf("JNI_SEQUENCE_OK (%d checks)\n", g_checks);
    return r;
}
