package ch.fhnw.woipv.nbody.simulation.hip;

import ch.fhnw.woipv.nbody.simulation.AbstractNBodySimulation;
import ch.fhnw.woipv.nbody.simulation.universe.UniverseGenerator;

import com.jogamp.opengl.GL3;

import java.nio.ByteBuffer;
import java.nio.ByteOrder;
import java.nio.FloatBuffer;

import static com.jogamp.opengl.GL.GL_ARRAY_BUFFER;

/**
 * Drop-in replacement for GPUBarnesHutNBodySimulation
 * (src/ch/fhnw/woipv/nbody/simulation/gpu/GPUBarnesHutNBodySimulation.java:38-448):
 * the same constructor, init / initGLBuffers / step / getNumberOfBodies, with
 * the OpenCL pipeline replaced by libnbody.so (HIP, MI355X) through JNI
 * (libnbody_jni.so: nbody_jni.c over nbody_jni_core.c). Java 8 compatible, like
 * the reference.
 *
 * The three-argument constructor runs plain Barnes-Hut with the reference's
 * compiled-in constants (calculateforce.cl:11-18). The four-argument one takes
 * every engine parameter (include/nbody.h, nb_params), including the learned
 * coarse-theta mode, whose network is trained with {@link #train} or installed
 * with {@link #setModel}.
 *
 * The device is headless: instead of sharing the VBOs with OpenCL
 * (clCreateFromGLBuffer, java:234-236), step() copies the float4 vertices
 * into direct buffers and uploads them with glBufferSubData.
 */
public class HipBarnesHutNBodySimulation extends AbstractNBodySimulation {

    static {
        System.loadLibrary("nbody_jni");
    }

    public static final int MODE_PLAIN = 0;
    public static final int MODE_LEARNED = 1;
    public static final int MODE_COARSE = 2;
    public static final int FP64_POSITIONS = 1;
    public static final int FUSED_UPDATE = 2;
    /** Floats in a model blob (NB_MODEL_FLOATS). */
    public static final int MODEL_FLOATS = 5955 + 2 * 24 + 2 * 3;

    /** Engine parameters (nb_params); the defaults are nb_default_params'. */
    public static final class Params {
        // reference constants (calculateforce.cl:11-18); theta and softening are the
        // square roots of the reference's THETA and EPSILON macros
        public float theta = (float) Math.sqrt(1.5);
        public float softening = 0.05f;
        public float dt = 0.025f;
        public int voteWidth = 16;
        public int flags = 0;
        public int device = 0;
        public float poolFactor = 2.0f;
        public int mode = MODE_PLAIN;
        public float coarseTheta = 0.95f;
        public float coarseAlpha = 0.02f;
        public float checkTol = 1e-3f;
        public int checkEvery = 0;
        public int checkSamples = 4096;
        public int checkAdapt = 0;

        float[] floats() {
            return new float[] {theta, softening, dt, poolFactor, coarseTheta, checkTol, coarseAlpha};
        }

        int[] ints() {
            return new int[] {voteWidth, flags, device, mode, checkEvery, checkSamples, checkAdapt};
        }
    }

    /** Training parameters (nb_train_params); the defaults are nb_default_train_params'. */
    public static final class TrainParams {
        public int steps = 50;
        public int groupsPerStep = 512;
        public int iters = 4000;
        public int batch = 4096;
        public float lr = 2e-3f;
        public float lrFinal = 1e-4f;
        public float beta1 = 0.9f;
        public float beta2 = 0.999f;
        public float adamEps = 1e-8f;
        public long seed = 1;
    }

    private final int nbodies;
    private final UniverseGenerator universeGenerator;
    private final Params params;
    private long handle;

    private GL3 gl;
    private int positionVBO = -1, velocityVBO = -1;
    private FloatBuffer positions, velocities;

    public HipBarnesHutNBodySimulation(final Mode mode, final int nbodies, final UniverseGenerator generator) {
        this(mode, nbodies, generator, new Params());
    }

    public HipBarnesHutNBodySimulation(final Mode mode, final int nbodies, final UniverseGenerator generator,
            final Params params) {
        super(mode);
        this.nbodies = nbodies;
        this.universeGenerator = generator;
        this.params = params;
    }

    @Override
    public void init(final GL3 gl) {
        this.handle = nativeCreate(nbodies, params.floats(), params.ints());
        final float[] x = new float[nbodies], y = new float[nbodies], z = new float[nbodies];
        final float[] vx = new float[nbodies], vy = new float[nbodies], vz = new float[nbodies];
        final float[] m = new float[nbodies];
        universeGenerator.generate(0, nbodies, x, y, z, vx, vy, vz, m);
        nativeSetState(handle, x, y, z, vx, vy, vz, m);
    }

    @Override
    public void initGLBuffers(final GL3 gl, final int positionVBO, final int velocityVBO) {
        this.gl = gl;
        this.positionVBO = positionVBO;
        this.velocityVBO = velocityVBO;
        positions = ByteBuffer.allocateDirect(16 * nbodies).order(ByteOrder.nativeOrder()).asFloatBuffer();
        velocities = ByteBuffer.allocateDirect(16 * nbodies).order(ByteOrder.nativeOrder()).asFloatBuffer();
    }

    @Override
    public void step() {
        nativeStep(handle, 1);
        if (mode == Mode.GL_INTEROP && positions != null) {
            nativeCopyVertices(handle, positions, velocities);
            if (gl != null) {
                gl.glBindBuffer(GL_ARRAY_BUFFER, positionVBO);
                gl.glBufferSubData(GL_ARRAY_BUFFER, 0, 16L * nbodies, positions);
                gl.glBindBuffer(GL_ARRAY_BUFFER, velocityVBO);
                gl.glBufferSubData(GL_ARRAY_BUFFER, 0, 16L * nbodies, velocities);
            }
        }
    }

    @Override
    public int getNumberOfBodies() {
        return nbodies;
    }

    /** Several steps in one call (one device error check at the end). */
    public void steps(final int k) {
        nativeStep(handle, k);
    }

    /** Body state by original index (the reference's printPosition / buffer reads, java:305-374). */
    public void getState(final float[] x, final float[] y, final float[] z, final float[] vx, final float[] vy,
            final float[] vz) {
        nativeGetState(handle, x, y, z, vx, vy, vz);
    }

    /** Trains the learned mode's network on pairs from this simulation and installs it;
     * returns {samples, loss_first, loss_last, gen_ms, train_ms}. */
    public double[] train(final TrainParams tp) {
        return nativeTrain(handle, new float[] {tp.lr, tp.lrFinal, tp.beta1, tp.beta2, tp.adamEps},
                new int[] {tp.steps, tp.groupsPerStep, tp.iters, tp.batch}, tp.seed);
    }

    public void setModel(final float[] model) {
        nativeSetModel(handle, model);
    }

    public float[] getModel() {
        return nativeGetModel(handle);
    }

    /** longs: step, cells, max_depth, interactions, error, fallbacks, live_cells,
     * cell_interactions, checks; doubles: stage_ms[0..5], last_check_median, check_ms,
     * coarse_alpha (nb_counters). */
    public void counters(final long[] l, final double[] d) {
        nativeCounters(handle, l, d);
    }

    /** Learned-mode control state (nb_control) for a checkpoint: longs steps_since_model,
     * check_seed, fell_back, fail_streak, fallbacks, checks; doubles coarse_alpha,
     * last_check_median. */
    public void getControl(final long[] l, final double[] d) {
        nativeGetControl(handle, l, d);
    }

    public void setControl(final long[] l, final double[] d) {
        nativeSetControl(handle, l, d);
    }

    /** {median, p90, max, mean} relative force error against the FP64 direct sum. */
    public double[] forceError(final int nsamples, final long seed) {
        return nativeForceError(handle, nsamples, seed);
    }

    public void dispose() {
        if (handle != 0) {
            nativeDestroy(handle);
            handle = 0;
        }
    }

    // native side: nbody_jni.c; failures throw IllegalArgumentException (bad
    // arguments) or IllegalStateException with nb_last_error()
    private static native long nativeCreate(int n, float[] fparams, int[] iparams);
    private static native void nativeSetState(long h, float[] x, float[] y, float[] z, float[] vx, float[] vy,
            float[] vz, float[] m);
    private static native void nativeGetState(long h, float[] x, float[] y, float[] z, float[] vx, float[] vy,
            float[] vz);
    private static native void nativeStep(long h, int steps);
    private static native void nativeCopyVertices(long h, FloatBuffer pos4, FloatBuffer vel4);
    private static native double[] nativeTrain(long h, float[] tf, int[] ti, long seed);
    private static native void nativeSetModel(long h, float[] model);
    private static native float[] nativeGetModel(long h);
    private static native void nativeCounters(long h, long[] l, double[] d);
    private static native void nativeGetControl(long h, long[] l, double[] d);
    private static native void nativeSetControl(long h, long[] l, double[] d);
    private static native double[] nativeForceError(long h, int nsamples, long seed);
    private static native void nativeDestroy(long h);
}
