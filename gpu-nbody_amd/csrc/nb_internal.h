// nb_internal.h -- device layout and kernel launch interface of the engine.
//
// HBM layout (all arrays are separate allocations, SoA):
//   bodies   x,y,z,vx,vy,vz,ax,ay,az,m : float[N]      (original body order)
//            x64,y64,z64               : double[N]     (NB_FP64_POSITIONS only)
//   octree   slot  : uint64[8*cap]   child slots, (epoch << 32) | child
//            ready : uint32[cap]     epoch at which summarize finished the cell
//            sready: uint32[cap]     epoch at which sort gave the cell its start
//            count : int32[cap]      bodies in the cell's subtree
//            start : int32[cap]      first sorted position of the subtree
//            com64 : double4[cap]    FP64 centre of mass (x,y,z) and mass (w)
//            cnode : float4[cap]     FP32 centre of mass and mass (traversal copy)
//            cchild: int32[8*cap]    compacted children for traversal: cells as
//                                    N + cell, bodies as their sorted position
//   sorted   sorted: int32[N]        body index at each sorted position
//            bnode : float4[N]       (x,y,z,m) of bodies in sorted order
//
// Node numbering follows the reference (buildtree.cl:277, sort.cl:223): a
// child value v < 0 is empty, v < N is a body, v >= N is cell v - N. Cell 0 is
// the root. Cells are allocated upward from 1, so every cell's index is larger
// than its parent's (the reference allocates downward from NN with the same
// property, buildtree.cl:109).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

constexpr int kMaxDepth = 64;       // reference MAXDEPTH, calculateforce.cl:12
constexpr int kBuildLevelLimit = 64; // cells live on levels < 64, bodies at depth <= 64
constexpr int kBBoxBlocks = 1024;   // partial-result slots of the bounding box
constexpr uint32_t kEmpty = 0xffffffffu;

enum DevError : int {
    kErrNone = 0,
    kErrPool = 1,  // cell pool exhausted
    kErrDepth = 2, // tree too deep
    kErrSync = 3,  // a bounded wait ran out
    kErrStack = 4, // traversal stack overflow
};

// Device-resident scalars of one simulation (the reference keeps these in
// one-element buffers: _step, _blockCount, _radius, _maxDepth, _bottom, _error).
struct Ctl {
    int step;            // reference _step, starts at -1 (java:167)
    uint32_t epoch;      // tree generation: tags child slots and ready flags
    int n_cells;         // cell allocation counter (reference _bottom)
    int max_depth;       // running maximum (reference _maxDepth, java:170)
    int error;           // DevError
    uint32_t bbox_count; // last-block detection (reference _blockCount)
    float radius;        // reference _radius
    float root[3];
    float dq[kMaxDepth]; // opening thresholds (calculateforce.cl:52-67)
    unsigned long long interactions; // cumulative body-node interactions
    int spin_limit;      // bound on the polling of summarize/sort
    int live_cells;      // cells linked into the tree by the last build
    int pad[2];
};

struct Dev {
    int n;
    int cap; // cells in the pool
    float theta2, eps2, dt;
    int fp64;
    Ctl *ctl;
    float *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m;
    double *x64, *y64, *z64;
    unsigned long long *slot;
    uint32_t *ready, *sready;
    int *count, *start;
    double4 *com64;
    float4 *cnode;
    int *cchild;
    int *sorted;
    float4 *bnode;
    float *bbox_part; // 6 * kBBoxBlocks
};

// ---- launchers (nb_tree.hip, nb_force.hip) ---------------------------------
hipError_t launch_bbox(const Dev &d, int blocks, hipStream_t st);
hipError_t launch_build(const Dev &d, int blocks, hipStream_t st);
hipError_t launch_summarize(const Dev &d, int blocks, hipStream_t st);
hipError_t launch_sort(const Dev &d, int blocks, hipStream_t st);
hipError_t launch_force(const Dev &d, int vote, int fused, hipStream_t st);
hipError_t launch_integrate(const Dev &d, hipStream_t st);
hipError_t launch_copy_vertices(const Dev &d, float4 *pos, float4 *vel, hipStream_t st);
hipError_t launch_paths(const Dev &d, unsigned long long *paths3, int *depth, hipStream_t st);
hipError_t launch_init_state(const Dev &d, hipStream_t st);

// persistent-grid sizes (all blocks must be co-resident for summarize/sort)
int summarize_blocks_per_cu(int device);
int sort_blocks_per_cu(int device);

} // namespace nb
