/*
 * nbody_jni.c -- JNI glue between HipBarnesHutNBodySimulation.java and the C
 * ABI of include/nbody.h. Build (needs a JDK for jni.h):
 *   gcc -O2 -shared -fPIC -I$JAVA_HOME/include -I$JAVA_HOME/include/linux \
 *       -I../../include nbody_jni.c -L.. -lnbody -Wl,-rpath,'$ORIGIN/..' -o libnbody_jni.so
 * Every failing call throws IllegalStateException(nb_last_error()), where the
 * reference let OpenCL exceptions propagate (GPUBarnesHutNBodySimulation.java:40-42).
 */
#include <jni.h>
#include <stdint.h>

#include "nbody.h"

#define CLS(name) Java_ch_fhnw_woipv_nbody_simulation_hip_HipBarnesHutNBodySimulation_##name

static void throw_state(JNIEnv *env, const nb_sim *sim, int code)
{
    jclass ex = (*env)->FindClass(env, "java/lang/IllegalStateException");
    const char *msg = nb_last_error(sim);
    (*env)->ThrowNew(env, ex, (msg && msg[0]) ? msg : nb_status_string(code));
}

JNIEXPORT jlong JNICALL CLS(nativeCreate)(JNIEnv *env, jclass cls, jint n, jfloat theta, jfloat softening,
                                          jfloat dt, jint vote, jint flags)
{
    (void)cls;
    nb_params p;
    nb_default_params(&p, n);
    p.theta = theta;
    p.softening = softening;
    p.dt = dt;
    p.vote_width = vote;
    p.flags = (uint32_t)flags;
    nb_sim *sim = NULL;
    int r = nb_create(&p, &sim);
    if (r != NB_OK) {
        throw_state(env, NULL, r);
        return 0;
    }
    return (jlong)(intptr_t)sim;
}

JNIEXPORT void JNICALL CLS(nativeSetState)(JNIEnv *env, jclass cls, jlong h, jfloatArray x, jfloatArray y,
                                           jfloatArray z, jfloatArray vx, jfloatArray vy, jfloatArray vz,
                                           jfloatArray m)
{
    (void)cls;
    nb_sim *sim = (nb_sim *)(intptr_t)h;
    jfloatArray arr[7] = {x, y, z, vx, vy, vz, m};
    float *p[7];
    for (int i = 0; i < 7; ++i) p[i] = (*env)->GetFloatArrayElements(env, arr[i], NULL);
    int r = nb_set_state(sim, p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    for (int i = 0; i < 7; ++i) (*env)->ReleaseFloatArrayElements(env, arr[i], p[i], JNI_ABORT);
    if (r != NB_OK) throw_state(env, sim, r);
}

JNIEXPORT void JNICALL CLS(nativeStep)(JNIEnv *env, jclass cls, jlong h, jint steps)
{
    (void)cls;
    nb_sim *sim = (nb_sim *)(intptr_t)h;
    int r = nb_step(sim, steps);
    if (r != NB_OK) throw_state(env, sim, r);
}

JNIEXPORT void JNICALL CLS(nativeCopyVertices)(JNIEnv *env, jclass cls, jlong h, jobject pos4, jobject vel4)
{
    (void)cls;
    nb_sim *sim = (nb_sim *)(intptr_t)h;
    float *p = pos4 ? (float *)(*env)->GetDirectBufferAddress(env, pos4) : NULL;
    float *v = vel4 ? (float *)(*env)->GetDirectBufferAddress(env, vel4) : NULL;
    int r = nb_copy_vertices(sim, p, v);
    if (r != NB_OK) throw_state(env, sim, r);
}

JNIEXPORT void JNICALL CLS(nativeDestroy)(JNIEnv *env, jclass cls, jlong h)
{
    (void)env;
    (void)cls;
    nb_destroy((nb_sim *)(intptr_t)h);
}
