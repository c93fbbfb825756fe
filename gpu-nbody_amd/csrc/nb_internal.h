// nb_internal.h -- device layout and kernel launch interface of the engine.
//
// HBM layout (separate allocations):
//   bodies   pm4   : float4[N]   (x, y, z, m), original body order
//            vel4  : float4[N]   (vx, vy, vz, 0)
//            acc4  : float4[N]   (ax, ay, az, 0)
//            pos64 : double4[N]  FP64 positions (NB_FP64_POSITIONS only)
//   octree   slot  : uint64[8*cap]  child slots, (epoch << 32) | child
//            cinfo : int4[cap]      (parent, level | octant << 8, epoch, 0), written when
//                                   the cell is published
//            count : int32[cap]     bodies in the cell's subtree
//            com64 : double4[cap]   FP64 centre of mass (x, y, z) and mass (w)
//            cchild: int32[8*cap]   children compacted to the front; cells as
//                                   N + cell, bodies by index (-1 after the last)
//            cdata : float4[8*cap]  (x, y, z, m) of those children: body position
//                                   and mass, or the child cell's FP32 centre of
//                                   mass and mass -- a popped cell is one
//                                   contiguous 160-byte read in the traversal;
//                                   slots past the last child are not written
//            cq64  : double[8*cap]  second moments about the centre of mass, learned
//                                   mode only: traceless Q (xx, yy, zz, xy, xz, yz),
//                                   trace T, 0
//            cqa/cqb: float4[8*cap] each child cell's (Qxx, Qyy, Qzz, T eps^2) and
//                                   (Qxy, Qxz, Qyz, 0), next to cdata; with cmac
//                                   (float2[8*cap], the opening-test constants)
//                                   written for cell children only
//   levels   lvl_list : int32[cap]  live cells grouped by level
//   sorted   bnode  : float4[N]     (x, y, z, m) in sorted order as the fused update
//                                   wrote them (the next build's positions)
//            sorted : int32[N]      body index at each sorted position: the
//                                   build's sorted kb_val[0], whose buffer the
//                                   sort stage swaps with this one
//   build    kb_key, kb_val : uint64[N] and int32[N], two of each: octant-path
//                                   keys and body indices through the radix
//                                   passes, the sorted pair in [0]
//            kb_deep: uint64[3N]    levels 21..83 of the bodies of equal-key runs
//            binary radix tree, int32[N] each: kb_ipar, kb_lpar (parents of
//                                   internal nodes and of sorted bodies), kb_lo
//                                   (first sorted body), kb_own / kb_ktop /
//                                   kb_obase (octree levels and first cell held)
//            kb_scan, kb_tdep: per scan tile of the cell counts, its sum and depth
//            kb_spl, kb_bcnt, kb_bfill, kb_bstart: the buckets of a warm build
//
// Body records are 16-byte AoS: streaming kernels still read them with fully
// coalesced 16-byte lanes, and the gathers of the tree stages touch one cache
// line per body instead of four.
//
// Node numbering follows the reference (buildtree.cl:277, sort.cl:223): a
// child value v < 0 is empty, v < N is a body, v >= N is cell v - N. Cell 0 is
// the root. Cells are numbered upward from 1 in octant-path order of their
// bodies (the reference takes them downward from NN, buildtree.cl:109); cell
// numbers carry no meaning beyond that.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

constexpr int kMaxDepth = 64;        // reference MAXDEPTH, calculateforce.cl:12
constexpr int kBuildLevelLimit = 64; // cells live on levels < 64, bodies at depth <= 64
constexpr int kBBoxBlocks = 1024;    // partial-result slots of the bounding box
constexpr int kLevelBlocks = 512;    // blocks of the level histogram / scatter
constexpr int kTailThreads = 512;    // the single block of the deepest summarize levels
constexpr unsigned char kNotLive = 0xff; // clvl of a cell that is not in the tree
constexpr int kCostKeys = 32;        // cost classes of the coarse traversal's launch order
constexpr uint32_t kEmpty = 0xffffffffu;

enum DevError : int {
    kErrNone = 0,
    kErrPool = 1,  // cell pool exhausted
    kErrDepth = 2, // tree too deep
    kErrSync = 3,  // inconsistent tree detected
    kErrStack = 4, // traversal stack overflow
};

// Device-resident scalars of one simulation (the reference keeps these in
// one-element buffers: _step, _radius, _maxDepth, _bottom, _error; its
// _blockCount, the bounding box's last-block counter, has no counterpart here:
// a second launch finishes the box).
struct Ctl {
    int step;            // reference _step, starts at -1 (java:167)
    uint32_t epoch;      // tree generation: tags child slots
    int n_cells;         // cell allocation counter (reference _bottom)
    int max_depth;       // running maximum (reference _maxDepth, java:170)
    int error;           // DevError
    float radius;        // reference _radius
    float root[3];
    float dq[kMaxDepth]; // opening thresholds (calculateforce.cl:52-67)
    unsigned long long interactions; // cumulative body-node interactions
    int levels;          // number of non-empty cell levels of the last tree
    int live_cells;      // cells linked into the tree by the last build
    unsigned long long cell_interactions; // cumulative cell interactions of the coarse traversal
    int lvl_start[kMaxDepth + 2]; // first lvl_list entry of each level
    int large_buckets;   // buckets of warm builds sorted through device memory, cumulative
};

struct Dev {
    int n;
    int cap; // cells in the pool
    float theta2, eps2, dt;
    float it2_coarse;    // learned mode: 1 / coarse_theta^2 (the extent test)
    float alpha2_coarse; // learned mode: coarse_alpha^2 (the relative test; 0 = off)
    int fp64;
    int quad;            // summarize also computes quadrupole moments
    Ctl *ctl;
    float4 *pm4, *vel4, *acc4;
    double4 *pos64;
    unsigned long long *slot;
    int4 *cinfo;
    int *count;
    double4 *com64;
    int *cchild;
    float4 *cdata;
    double *cq64;
    float4 *cqa;
    float4 *cqb;
    float2 *cmac;       // per child slot: (b^2 / theta_c^2 + eps^2, (m b^3)^(2/5)) of the child's extent b
    int *lvl_list;
    unsigned char *clvl; // cap: each cell's level, or kNotLive (level histogram -> scatter)
    int *lvl_blockhist; // kLevelBlocks * kMaxDepth
    float4 *bnode;
    int *sorted;
    float4 *probe;      // accelerations of a probe force pass (error checks)
    float *bbox_part; // 6 * kBBoxBlocks
    int *gcost;       // per 64-body group: cost of its last coarse traversal (children tested)
    int *gorder;      // the groups, most expensive first (the coarse traversal's launch order)
    // the key-sorted build (nb_keys.hip): octant-path keys and bodies (two
    // buffers each for the radix passes; the sorted result ends in [0]), the
    // deep key words of runs (3 * n), the radix sort's per-tile digit counts
    // and offsets (256 * tiles) and digit totals (256), and the binary radix
    // tree over the sorted keys (n each)
    unsigned long long *kb_key[2];
    int *kb_val[2];
    unsigned long long *kb_deep;
    unsigned *kb_hist, *kb_offs, *kb_tot;
    int *kb_ipar, *kb_lpar, *kb_lo, *kb_own, *kb_ktop, *kb_obase, *kb_scan, *kb_tdep;
    // warm builds: splitters (S = bucket_samples(n)), bucket counts, fill
    // counters (S + 1 each) and starts (S + 2)
    unsigned long long *kb_spl;
    unsigned *kb_bcnt, *kb_bfill, *kb_bstart;
    int *kb_big; // S + 2: the count of buckets above kBucketCap, then their numbers
};

// ---- launchers (nb_tree.hip, nb_force.hip) ---------------------------------
hipError_t launch_bbox(const Dev &d, int blocks, hipStream_t st);
// the build from sorted octant-path keys (the same child slots and cell records)
// warm: start from sorted[] and the last sorted keys; staged: bnode[] holds
// the records in sorted[] order
hipError_t launch_build_keys(const Dev &d, bool warm, bool staged, hipStream_t st);
constexpr int kKeyTile = 2048;     // keys per radix-sort tile (first build)
constexpr int kKeyScanTile = 1024; // cell counts per scan tile
constexpr int kBucketT = 256;      // warm builds: one splitter per kBucketT bodies
constexpr int kBucketCap = 512;    // buckets sorted by one wave up to this size
inline int bucket_samples(int n) { return (n + kBucketT - 1) / kBucketT; }
hipError_t launch_levels(const Dev &d, hipStream_t st); // group live cells by level
// Summarize and sort read each level's cell count on the device; the host's
// cells_hint (the level's size in an earlier tree) only sizes the grid. The
// block launches take levels lo .. hi-1 in one block (hi = kMaxDepth + 1:
// every level from lo down, however deep the tree): the deepest levels and
// the top ones, where the levels hold few cells.
hipError_t launch_summarize_level(const Dev &d, int level, int cells_hint, hipStream_t st);
hipError_t launch_summarize_block(const Dev &d, int lo, int hi, hipStream_t st);
// mode: 0 = store accelerations (reference semantics), 1 = also integrate
// (fused kick-drift-kick), 2 = probe: write accelerations to d.probe only
hipError_t launch_force(const Dev &d, int vote, int mode, hipStream_t st);
// FP64 direct-sum accelerations of the bodies idx[0..ns) (G = 1), into out;
// out[i].w = sum_{j != idx[i]} m_j / sqrt(r^2 + eps^2) (the softened potential, negated)
// when phi, else 0
hipError_t launch_direct64(const Dev &d, const int *idx, int ns, double4 *part, double4 *out, bool phi,
                           hipStream_t st);
hipError_t launch_integrate(const Dev &d, hipStream_t st);
// FP64 per-block sums of m, m v, m v^2 / 2, m x cross v, m x (11 doubles per block)
constexpr int kMomentBlocks = 512;
hipError_t launch_moments(const Dev &d, double *part, hipStream_t st);
// bodies first .. first + count - 1 into pos[0 .. count) / vel[0 .. count) (either may be null)
hipError_t launch_copy_vertices(const Dev &d, int first, int count, float4 *pos, float4 *vel, hipStream_t st);
hipError_t launch_paths(const Dev &d, unsigned long long *paths3, int *depth, hipStream_t st);

// ---- learned coarse-theta correction (nb_learned.hip) -----------------------
constexpr int kFeat = 24, kHid = 64, kOut = 3;
constexpr int kMlpParams = kFeat * kHid + kHid + kHid * kHid + kHid + kHid * kOut + kOut; // 5955
// parameter offsets: W1[24][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]
constexpr int kOffB1 = kFeat * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut;

// Inference network for the matrix cores, in the transposed form
// Z^T = W^T X^T (bodies are the columns), stored as the per-lane A-operand
// fragments of v_mfma_f32_16x16x32_bf16 (8 BF16 per lane, lane l: row l % 16
// of the 16-row block, k group g = l / 16):
//   w1f[rb][lane][i]      W1[8 g + i][16 rb + l % 16]            (inputs 24..31: 0)
//   w2f[rb][kb][lane][i]  W2[pk(kb, g, i)][16 rb + l % 16]
//   w3f[kb][lane][i]      W3[pk(kb, g, i)][l % 16] * sd_y         (outputs 3..15: 0)
// where pk(kb, g, i) = 32 kb + 4 g + i for i < 4 and 32 kb + 16 + 4 g + i - 4
// for i >= 4 is the order in which a lane already holds the previous layer's
// outputs, so activations pass from layer to layer without moving between
// lanes. FP32 biases (b3 with the target mean folded in, 16 entries) and the
// feature normalization (mean, 1/sd) follow.
struct Mlp {
    const uint16_t *w1f, *w2f, *w3f;
    const float *b1, *b2, *b3;
    const float *xmu, *xrs;
};
constexpr int kW1Frag = 4 * 64 * 8, kW2Frag = 4 * 2 * 64 * 8, kW3Frag = 2 * 64 * 8; // bf16 counts
// normalized features and training targets are clipped to [-kClipZ, kClipZ]:
// both have rare tails of tens of standard deviations (bodies whose coarse
// acceleration nearly cancels), which would otherwise dominate the squared loss
constexpr float kClipZ = 8.0f;
constexpr int kGradChunks = 16; // row chunks of the weight-gradient reduction
constexpr size_t kMlpBufBytes = (size_t)(kW1Frag + kW2Frag + kW3Frag) * 2 + (64 + 64 + 16 + 24 + 24) * 4;

struct Learn {
    Mlp mlp;
};

// What the coarse traversal does with its result: the learned step (features
// through the network, corrected acceleration, body update), raw FP32
// features and frames of listed groups (training pairs), accelerations of
// listed groups (training targets), or the body update with the coarse
// acceleration alone.
// kOutAccel opens cells with the force stage's table `dqtab` (the fine
// traversal of the training targets); the others with the extent and
// relative tests of the coarse traversal (dqtab unused).
enum CoarseOut { kOutLearned = 0, kOutRaw = 1, kOutAccel = 2, kOutFinish = 3 };
hipError_t launch_coarse(const Dev &d, const Learn &L, int out, const float *dqtab, const int *groups,
                         int ngroups, float *xraw, float4 *frame_out, float4 *acc_out, int fmode,
                         hipStream_t st);
// The network alone on raw FP32 features x[rows][24] -> residual y[rows][3]
// (the same matrix-core code as the learned step)
hipError_t launch_mlp_rows(const Learn &L, const float *xraw, int rows, float *y, hipStream_t st);
// training targets (a_fine - a) in the frame over |a|
hipError_t launch_make_target(const float4 *frame, const float4 *afine, int rows, float *t, hipStream_t st);
// per-column mean and standard deviation (sd 1 for a constant column), then
// x = (x - mean) / sd in place
hipError_t launch_standardize(float *x, int rows, int cols, float *mu, float *sd, hipStream_t st);
hipError_t launch_gather_acc(const Dev &d, const int *idx, int ns, float4 *out, hipStream_t st);
// d.gorder = the (n + 63) / 64 groups by the cost of their last coarse
// traversal (d.gcost), most expensive first, so that the longest waves start
// first and the launch does not end on a few of them
hipError_t launch_group_order(const Dev &d, hipStream_t st);

struct Train {
    float *P, *G, *M, *V;  // parameters, gradient, Adam moments (kMlpParams each)
    const float *X, *T;    // dataset: normalized features [S][24], targets [S][3]
    int S;
    float *Z1, *H1, *Z2, *H2, *dZ1, *dZ2, *dY; // [B][64] (dY [B][3])
    float *gpart;          // weight-gradient partial sums: kGradChunks x (kHid + 1) x kHid
    float *loss_part;      // one per 256 rows of the batch
    float *loss_hist;      // one per iteration
};
// forward + backward of one minibatch (rows drawn with `seed`, or rows 0..B-1
// when seed is 0) into t.G and t.loss_hist[it]; then, if adam, one Adam step
hipError_t launch_train_iter(const Train &t, int B, int it, uint64_t seed, bool adam, float lr, float beta1,
                             float beta2, float eps, float c1, float c2, hipStream_t st);

} // namespace nb
