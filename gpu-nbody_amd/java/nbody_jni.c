/*
 * nbody_jni.c -- JNI glue between HipBarnesHutNBodySimulation.java and the C
 * ABI of include/nbody.h. It moves data between the JVM and nbody_jni_core.c,
 * which does the checks and the calls. Build (needs a JDK for jni.h):
 *   gcc -O2 -shared -fPIC -I$JAVA_HOME/include -I$JAVA_HOME/include/linux \
 *       -Iinclude -Igpu-nbody_amd/java gpu-nbody_amd/java/nbody_jni.c \
 *       gpu-nbody_amd/java/nbody_jni_core.c -Lgpu-nbody_amd -lnbody \
 *       -Wl,-rpath,'$ORIGIN' -o gpu-nbody_amd/libnbody_jni.so
 * A failing call throws IllegalArgumentException (NB_E_ARG) or
 * IllegalStateException (every other code) with the library's message, where
 * the reference let OpenCL exceptions propagate
 * (GPUBarnesHutNBodySimulation.java:40-42).
 */
#include <jni.h>
#include <stdint.h>

#include "nbody_jni_core.h"

#define CLS(name) Java_ch_fhnw_woipv_nbody_simulation_hip_HipBarnesHutNBodySimulation_##name

static void throw_for(JNIEnv *env, const nb_sim *sim, int code, const char *why)
{
    const char *cls = code == NB_E_ARG ? "java/lang/IllegalArgumentException" : "java/lang/IllegalStateException";
    jclass ex = (*env)->FindClass(env, cls);
    if (ex) (*env)->ThrowNew(env, ex, nbj_message(sim, code, why));
}

#define SIM(h) ((nb_sim *)(intptr_t)(h))

/* Small fixed-size arrays are copied (Get*ArrayRegion); a null or wrongly
 * sized one reaches the core as NULL / its length, which rejects it. */
static jint get_floats(JNIEnv *env, jfloatArray a, float *out, jint cap)
{
    if (!a) return -1;
    const jint len = (*env)->GetArrayLength(env, a);
    if (len > cap) return len;
    (*env)->GetFloatArrayRegion(env, a, 0, len, out);
    return len;
}

static jint get_ints(JNIEnv *env, jintArray a, int32_t *out, jint cap)
{
    if (!a) return -1;
    const jint len = (*env)->GetArrayLength(env, a);
    if (len > cap) return len;
    (*env)->GetIntArrayRegion(env, a, 0, len, (jint *)out);
    return len;
}

JNIEXPORT jlong JNICALL CLS(nativeCreate)(JNIEnv *env, jclass cls, jint n, jfloatArray fparams, jintArray iparams)
{
    (void)cls;
    float f[NBJ_NF];
    int32_t i[NBJ_NI];
    const jint nf = get_floats(env, fparams, f, NBJ_NF), ni = get_ints(env, iparams, i, NBJ_NI);
    nb_sim *sim = NULL;
    const char *why = NULL;
    int r = nbj_create(n, nf >= 0 ? f : NULL, nf, ni >= 0 ? i : NULL, ni, &sim, &why);
    if (r != NB_OK) {
        throw_for(env, NULL, r, why);
        return 0;
    }
    return (jlong)(intptr_t)sim;
}

/* Pins up to 7 float[] (the body arrays); NULL entries stay NULL. */
static int pin(JNIEnv *env, jfloatArray *arr, int k, float **p, int32_t *len)
{
    for (int j = 0; j < k; ++j) {
        p[j] = NULL;
        len[j] = 0;
    }
    for (int j = 0; j < k; ++j) {
        if (!arr[j]) continue;
        len[j] = (*env)->GetArrayLength(env, arr[j]);
        p[j] = (*env)->GetFloatArrayElements(env, arr[j], NULL);
        if (!p[j]) return -1; /* OutOfMemoryError is pending */
    }
    return 0;
}

static void unpin(JNIEnv *env, jfloatArray *arr, int k, float **p, jint mode)
{
    for (int j = 0; j < k; ++j)
        if (p[j]) (*env)->ReleaseFloatArrayElements(env, arr[j], p[j], mode);
}

JNIEXPORT void JNICALL CLS(nativeSetState)(JNIEnv *env, jclass cls, jlong h, jfloatArray x, jfloatArray y,
                                           jfloatArray z, jfloatArray vx, jfloatArray vy, jfloatArray vz,
                                           jfloatArray m)
{
    (void)cls;
    jfloatArray arr[7] = {x, y, z, vx, vy, vz, m};
    float *p[7];
    int32_t len[7];
    if (pin(env, arr, 7, p, len) == 0) {
        const char *why = NULL;
        int r = nbj_set_state(SIM(h), p, len, &why);
        unpin(env, arr, 7, p, JNI_ABORT); /* read only: nothing to copy back */
        if (r != NB_OK) throw_for(env, SIM(h), r, why);
    } else {
        unpin(env, arr, 7, p, JNI_ABORT);
    }
}

JNIEXPORT void JNICALL CLS(nativeGetState)(JNIEnv *env, jclass cls, jlong h, jfloatArray x, jfloatArray y,
                                           jfloatArray z, jfloatArray vx, jfloatArray vy, jfloatArray vz)
{
    (void)cls;
    jfloatArray arr[6] = {x, y, z, vx, vy, vz};
    float *p[6];
    int32_t len[6];
    if (pin(env, arr, 6, p, len) == 0) {
        const char *why = NULL;
        int r = nbj_get_state(SIM(h), p, len, &why);
        unpin(env, arr, 6, p, r == NB_OK ? 0 : JNI_ABORT);
        if (r != NB_OK) throw_for(env, SIM(h), r, why);
    } else {
        unpin(env, arr, 6, p, JNI_ABORT);
    }
}

JNIEXPORT void JNICALL CLS(nativeStep)(JNIEnv *env, jclass cls, jlong h, jint steps)
{
    (void)cls;
    const char *why = NULL;
    int r = nbj_step(SIM(h), steps, &why);
    if (r != NB_OK) throw_for(env, SIM(h), r, why);
}

JNIEXPORT void JNICALL CLS(nativeCopyVertices)(JNIEnv *env, jclass cls, jlong h, jobject pos4, jobject vel4)
{
    (void)cls;
    float *p = pos4 ? (float *)(*env)->GetDirectBufferAddress(env, pos4) : NULL;
    float *v = vel4 ? (float *)(*env)->GetDirectBufferAddress(env, vel4) : NULL;
    if ((pos4 && !p) || (vel4 && !v)) {
        throw_for(env, SIM(h), NB_E_ARG, "vertex buffers must be direct buffers");
        return;
    }
    /* capacity of a direct FloatBuffer, in floats */
    const jlong pc = pos4 ? (*env)->GetDirectBufferCapacity(env, pos4) : 0;
    const jlong vc = vel4 ? (*env)->GetDirectBufferCapacity(env, vel4) : 0;
    const char *why = NULL;
    int r = nbj_copy_vertices(SIM(h), p, pc, v, vc, &why);
    if (r != NB_OK) throw_for(env, SIM(h), r, why);
}

JNIEXPORT jdoubleArray JNICALL CLS(nativeTrain)(JNIEnv *env, jclass cls, jlong h, jfloatArray tf, jintArray ti,
                                                jlong seed)
{
    (void)cls;
    float f[NBJ_TF];
    int32_t i[NBJ_TI];
    const jint nf = get_floats(env, tf, f, NBJ_TF), ni = get_ints(env, ti, i, NBJ_TI);
    double rep[NBJ_TREP];
    const char *why = NULL;
    int r = nbj_train(SIM(h), nf >= 0 ? f : NULL, nf, ni >= 0 ? i : NULL, ni, seed, rep, NBJ_TREP, &why);
    if (r != NB_OK) {
        throw_for(env, SIM(h), r, why);
        return NULL;
    }
    jdoubleArray out = (*env)->NewDoubleArray(env, NBJ_TREP);
    if (out) (*env)->SetDoubleArrayRegion(env, out, 0, NBJ_TREP, rep);
    return out;
}

JNIEXPORT void JNICALL CLS(nativeSetModel)(JNIEnv *env, jclass cls, jlong h, jfloatArray model)
{
    (void)cls;
    float buf[NB_MODEL_FLOATS];
    const jint len = get_floats(env, model, buf, NB_MODEL_FLOATS);
    const char *why = NULL;
    int r = nbj_set_model(SIM(h), len >= 0 ? buf : NULL, len, &why);
    if (r != NB_OK) throw_for(env, SIM(h), r, why);
}

JNIEXPORT jfloatArray JNICALL CLS(nativeGetModel)(JNIEnv *env, jclass cls, jlong h)
{
    (void)cls;
    float buf[NB_MODEL_FLOATS];
    const char *why = NULL;
    int r = nbj_get_model(SIM(h), buf, NB_MODEL_FLOATS, &why);
    if (r != NB_OK) {
        throw_for(env, SIM(h), r, why);
        return NULL;
    }
    jfloatArray out = (*env)->NewFloatArray(env, NB_MODEL_FLOATS);
    if (out) (*env)->SetFloatArrayRegion(env, out, 0, NB_MODEL_FLOATS, buf);
    return out;
}

JNIEXPORT void JNICALL CLS(nativeCounters)(JNIEnv *env, jclass cls, jlong h, jlongArray l, jdoubleArray d)
{
    (void)cls;
    int64_t lv[NBJ_CL];
    double dv[NBJ_CD];
    const char *why = NULL;
    if (!l || !d || (*env)->GetArrayLength(env, l) < NBJ_CL || (*env)->GetArrayLength(env, d) < NBJ_CD) {
        throw_for(env, SIM(h), NB_E_ARG, "counter arrays must hold 9 longs and 9 doubles");
        return;
    }
    int r = nbj_counters(SIM(h), lv, NBJ_CL, dv, NBJ_CD, &why);
    if (r != NB_OK) {
        throw_for(env, SIM(h), r, why);
        return;
    }
    (*env)->SetLongArrayRegion(env, l, 0, NBJ_CL, (const jlong *)lv);
    (*env)->SetDoubleArrayRegion(env, d, 0, NBJ_CD, dv);
}

JNIEXPORT void JNICALL CLS(nativeGetControl)(JNIEnv *env, jclass cls, jlong h, jlongArray l, jdoubleArray d)
{
    (void)cls;
    int64_t lv[NBJ_KL];
    double dv[NBJ_KD];
    const char *why = NULL;
    if (!l || !d || (*env)->GetArrayLength(env, l) < NBJ_KL || (*env)->GetArrayLength(env, d) < NBJ_KD) {
        throw_for(env, SIM(h), NB_E_ARG, "control arrays must hold 6 longs and 2 doubles");
        return;
    }
    int r = nbj_get_control(SIM(h), lv, NBJ_KL, dv, NBJ_KD, &why);
    if (r != NB_OK) {
        throw_for(env, SIM(h), r, why);
        return;
    }
    (*env)->SetLongArrayRegion(env, l, 0, NBJ_KL, (const jlong *)lv);
    (*env)->SetDoubleArrayRegion(env, d, 0, NBJ_KD, dv);
}

JNIEXPORT void JNICALL CLS(nativeSetControl)(JNIEnv *env, jclass cls, jlong h, jlongArray l, jdoubleArray d)
{
    (void)cls;
    int64_t lv[NBJ_KL];
    double dv[NBJ_KD];
    const char *why = NULL;
    if (!l || !d || (*env)->GetArrayLength(env, l) != NBJ_KL || (*env)->GetArrayLength(env, d) != NBJ_KD) {
        throw_for(env, SIM(h), NB_E_ARG, "control arrays must hold 6 longs and 2 doubles");
        return;
    }
    (*env)->GetLongArrayRegion(env, l, 0, NBJ_KL, (jlong *)lv);
    (*env)->GetDoubleArrayRegion(env, d, 0, NBJ_KD, dv);
    int r = nbj_set_control(SIM(h), lv, NBJ_KL, dv, NBJ_KD, &why);
    if (r != NB_OK) throw_for(env, SIM(h), r, why);
}

JNIEXPORT jdoubleArray JNICALL CLS(nativeForceError)(JNIEnv *env, jclass cls, jlong h, jint nsamples, jlong seed)
{
    (void)cls;
    double st[4];
    const char *why = NULL;
    int r = nbj_force_error(SIM(h), nsamples, seed, st, 4, &why);
    if (r != NB_OK) {
        throw_for(env, SIM(h), r, why);
        return NULL;
    }
    jdoubleArray out = (*env)->NewDoubleArray(env, 4);
    if (out) (*env)->SetDoubleArrayRegion(env, out, 0, 4, st);
    return out;
}

JNIEXPORT void JNICALL CLS(nativeDestroy)(JNIEnv *env, jclass cls, jlong h)
{
    (void)env;
    (void)cls;
    nbj_destroy(SIM(h));
}
