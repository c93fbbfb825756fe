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
 * (libnbody_jni.so, nbody_jni.c). Java 8 compatible, like the reference.
 *
 * The device is headless: instead of sharing the VBOs with OpenCL
 * (clCreateFromGLBuffer, java:234-236), step() copies the float4 vertices
 * into direct buffers and uploads them with glBufferSubData.
 */
public class HipBarnesHutNBodySimulation extends AbstractNBodySimulation {

    static {
        System.loadLibrary("nbody_jni");
    }

    // reference constants (calculateforce.cl:11-18); theta and softening are the
    // square roots of the reference's THETA and EPSILON macros
    public static final float THETA = (float) Math.sqrt(1.5);
    public static final float SOFTENING = 0.05f;
    public static final float TIMESTEP = 0.025f;
    public static final int VOTE_WIDTH = 16;

    private final int nbodies;
    private final UniverseGenerator universeGenerator;
    private long handle;

    private GL3 gl;
    private int positionVBO = -1, velocityVBO = -1;
    private FloatBuffer positions, velocities;

    public HipBarnesHutNBodySimulation(final Mode mode, final int nbodies, final UniverseGenerator generator) {
        super(mode);
        this.nbodies = nbodies;
        this.universeGenerator = generator;
    }

    @Override
    public void init(final GL3 gl) {
        this.handle = nativeCreate(nbodies, THETA, SOFTENING, TIMESTEP, VOTE_WIDTH, 0);
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

    public void dispose() {
        if (handle != 0) {
            nativeDestroy(handle);
            handle = 0;
        }
    }

    // native side: nbody_jni.c; failures throw IllegalStateException with nb_last_error()
    private static native long nativeCreate(int n, float theta, float softening, float dt, int vote, int flags);
    private static native void nativeSetState(long h, float[] x, float[] y, float[] z, float[] vx, float[] vy,
            float[] vz, float[] m);
    private static native void nativeStep(long h, int steps);
    private static native void nativeCopyVertices(long h, FloatBuffer pos4, FloatBuffer vel4);
    private static native void nativeDestroy(long h);
}
