// nb_engine.cpp -- host side of the engine and the C ABI of include/nbody.h.
//
// Replaces GPUBarnesHutNBodySimulation (the reference's host driver): device
// selection, node-pool sizing, buffer allocation, the per-step launch order
// and the error checks. One HIP stream per handle at default priority; a step
// is six launches with no host synchronization between them, and nb_step
// reads the device error word once at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nbody.h"
#include "nb_internal.h"

struct nb_sim {
    nb_params p;
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    nb::Dev d{};
    int bbox_blocks = 1;
    bool have_state = false;
    bool timing = false;
    std::vector<hipEvent_t> tev; // per-step stage events (timing mode)
    double stage_ms[8] = {};
    unsigned long long inter_before = 0;
    long long last_interactions = 0, last_cell_interactions = 0, cell_inter_before = 0;
    long long fallbacks = 0;
    double last_check_median = 0.0;
    int *chk_idx = nullptr;             // sampled-error check buffers, allocated on first use
    double4 *chk_part = nullptr, *chk_out = nullptr;
    int chk_cap = 0;
    // level sizes of an earlier tree (the last one read back): they size the
    // grids of the summarize and sort launches, which read the current sizes
    // on the device, so no step waits for them
    bool hint_valid = false;
    int hint_levels = 0;
    int hint_start[nb::kMaxDepth + 2] = {};
    int split = 0; // levels split.. go to one single-block launch (the deepest levels)
    int head = 0;  // levels ..head-1 go to another (the top levels)
    // learned mode
    nb::Learn L{};
    void *mlpbuf = nullptr;      // packed inference network (nb::kMlpBufBytes)
    std::vector<float> model;    // installed model blob (NB_MODEL_FLOATS)
    bool have_model = false;
    bool fell_back = false;      // a sampled check failed: plain theta until a new model
    float alpha_now = 0.f;       // relative-test alpha in effect (check_adapt tightens it)
    int fail_streak = 0;         // failed checks in a row
    long long steps_since_model = 0;
    uint64_t check_seed = 0x5eed;
    float4 *chk_acc = nullptr;
    hipEvent_t chk_ev[2] = {nullptr, nullptr}; // bracket the sampled check's direct sum
    double check_ms = 0.0;                     // device time of the sampled checks so far
    long long checks = 0;
    // the direct sum of a check step runs on a second stream (default
    // priority) beside the step's tree stages and traversal
    hipStream_t side = nullptr;
    hipEvent_t chk_go = nullptr; // the step's positions are final (recorded on stream)
    int32_t *chk_host = nullptr; // pinned sample indices, so the upload never blocks the host
    int chk_host_cap = 0, chk_ns = 0;
    // the coarse traversal's group order for the next pass is worked out on a
    // stream of its own (default priority) right after each pass, beside the
    // next step's tree stages instead of in front of its traversal
    hipStream_t ord = nullptr;
    hipEvent_t ord_in = nullptr, ord_out = nullptr;
    bool ord_ready = false; // d.gorder holds an order once ord_out has passed
    bool keys_warm = false; // d.sorted holds the last tree's body order (the build starts from it)
    bool staged = false;    // d.bnode holds pm4 in d.sorted order (the last fused update wrote it)
    bool sort_pending = false; // d.kb_val[0] holds the order of a build the sort stage has not taken yet
    // copyVertices staging: 2 N float4 (positions, then velocities), allocated
    // by the first nb_copy_vertices call and kept until nb_destroy
    float4 *vtx = nullptr;
    std::string err;
};

// Staging of nb_copy_vertices_noalloc: a file-scope device array (one copy on
// each GPU, held until the process ends), so that entry point allocates
// nothing. Bodies go through it kNoAllocChunk at a time: positions in the
// first half, velocities in the second (2 MiB in all).
constexpr int kNoAllocChunk = 65536;
static __device__ __align__(256) float4 scratch[2 * kNoAllocChunk];

namespace {

// Device address of `scratch` on each GPU, looked up once per device by the
// first nb_create on it (host code never passes the array's own name, which
// is a host-side stand-in, not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float4 *g_scratch_ptr[kMaxDevices] = {};

// message of the last nb_create that failed on this thread (nb_last_error(NULL))
thread_local std::string g_create_error;

int fail(nb_sim *s, int code, const char *fmt, ...)
{
    if (s) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        s->err = buf;
    }
    return code;
}

#define HIP_TRY(sim, expr)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(sim, NB_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));            \
    } while (0)

// Cell pool: the reference's NUMBER_OF_NODES = 2N (java:219-227) scaled by
// pool_factor / 2. The build numbers the tree's cells densely from 0, so the
// pool holds exactly the cells a tree may have.
long long pool_cells(const nb_params *p)
{
    return (long long)std::max(16384.0, std::ceil((double)p->pool_factor * p->n_bodies));
}

int validate(const nb_params *p)
{
    if (!p) return NB_E_ARG;
    if (p->n_bodies < 1) return NB_E_ARG;
    if (!(p->theta >= 0.0f) || !std::isfinite(p->theta)) return NB_E_ARG;
    if (!(p->softening > 0.0f) || !std::isfinite(p->softening)) return NB_E_ARG;
    if (!std::isfinite(p->dt)) return NB_E_ARG;
    if (p->vote_width != 16 && p->vote_width != 32 && p->vote_width != 64) return NB_E_ARG;
    if (p->flags & ~(NB_FP64_POSITIONS | NB_FUSED_UPDATE)) return NB_E_ARG;
    if (!(p->pool_factor >= 1.0f) || p->pool_factor > 64.0f) return NB_E_ARG;
    if (p->mode != NB_MODE_PLAIN && p->mode != NB_MODE_LEARNED && p->mode != NB_MODE_COARSE) return NB_E_ARG;
    if (!(p->coarse_theta > 0.0f) || !std::isfinite(p->coarse_theta)) return NB_E_ARG;
    if (!(p->coarse_alpha >= 0.0f) || !std::isfinite(p->coarse_alpha)) return NB_E_ARG;
    // the coarse traversal is wave-uniform: one vote group is one wave64; its
    // extent test |d| >= b / theta_c keeps a body out of every cell it accepts
    // only for theta_c < 1
    if (p->mode != NB_MODE_PLAIN &&
        (p->vote_width != 64 || !(p->coarse_theta >= p->theta) || !(p->coarse_theta < 1.0f)))
        return NB_E_ARG;
    if (!(p->check_tol >= 0.0f) || p->check_every < 0 || p->check_samples < 0) return NB_E_ARG;
    if (p->check_adapt != 0 && p->check_adapt != 1) return NB_E_ARG;
    // the traversal stack packs (cell << 6 | depth) into 32 bits
    if ((double)pool_cells(p) > (double)(1 << 26)) return NB_E_ARG;
    return NB_OK;
}

template <class T> hipError_t dalloc(T **p, size_t count)
{
    return hipMalloc((void **)p, count * sizeof(T) + 64);
}

nb::Ctl initial_ctl(const nb_sim *s, uint32_t epoch)
{
    nb::Ctl c{};
    c.step = -1;     // GPUBarnesHutNBodySimulation.java:167
    c.epoch = epoch;
    c.n_cells = 1;
    c.max_depth = 1; // GPUBarnesHutNBodySimulation.java:170
    c.error = 0;
    c.interactions = 0;
    c.cell_interactions = 0;
    (void)s;
    return c;
}

int read_ctl(nb_sim *s, nb::Ctl *out)
{
    HIP_TRY(s, hipMemcpyAsync(out, s->d.ctl, sizeof(nb::Ctl), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}

// Keeps the level sizes of a tree read back to size later launches.
void update_hint(nb_sim *s, const nb::Ctl &c)
{
    if (c.error || c.levels <= 0) return;
    s->hint_valid = true;
    s->hint_levels = c.levels;
    std::memcpy(s->hint_start, c.lvl_start, sizeof s->hint_start);
}

int check_device_error(nb_sim *s)
{
    nb::Ctl c;
    int r = read_ctl(s, &c);
    if (r) return r;
    update_hint(s, c);
    s->last_interactions = (long long)(c.interactions - s->inter_before);
    s->inter_before = c.interactions;
    s->last_cell_interactions = (long long)(c.cell_interactions - s->cell_inter_before);
    s->cell_inter_before = c.cell_interactions;
    switch (c.error) {
    case nb::kErrNone: return NB_OK;
    case nb::kErrPool: return fail(s, NB_E_POOL, "octree cell pool exhausted (%d cells)", s->d.cap);
    case nb::kErrDepth: return fail(s, NB_E_DEPTH, "octree deeper than %d levels", NB_MAX_DEPTH);
    case nb::kErrSync: return fail(s, NB_E_SYNC, "bounded device wait ran out (inconsistent tree)");
    case nb::kErrStack: return fail(s, NB_E_STACK, "force traversal stack overflow");
    default: return fail(s, NB_E_HIP, "unknown device error %d", c.error);
    }
}

int hint_cells(const nb_sim *s, int level)
{
    return level < s->hint_levels ? s->hint_start[level + 1] - s->hint_start[level] : 0;
}

// Levels whose expected size is at most this many cells (two passes of the
// block each) go to the single-block launches: from the bottom up to the
// first larger level, and from the top down to the first larger level.
constexpr int kTailCells = 2 * nb::kTailThreads / 8;

// Groups the live cells by level. Without level sizes of an earlier tree (the
// first step of a state) it reads this tree's back, the only host wait inside
// a step; later steps size their launches from the last sizes read.
int prepare_levels(nb_sim *s)
{
    hipError_t e = nb::launch_levels(s->d, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "level grouping launch: %s", hipGetErrorString(e));
    if (!s->hint_valid) {
        nb::Ctl c;
        int r = read_ctl(s, &c);
        if (r) return r;
        update_hint(s, c);
    }
    int split = s->hint_levels;
    while (split > 0 && hint_cells(s, split - 1) <= kTailCells) --split;
    int head = 0;
    while (head < split && hint_cells(s, head) <= kTailCells) ++head;
    s->split = split;
    s->head = head;
    return NB_OK;
}

// Device address of the force stage's opening table inside Ctl.
const float *dq_plain_ptr(const nb_sim *s)
{
    return reinterpret_cast<const float *>(reinterpret_cast<const char *>(s->d.ctl) + offsetof(nb::Ctl, dq));
}

// The force method in effect: the configured mode, or plain theta after a
// failed sampled check.
int force_method(const nb_sim *s) { return s->fell_back ? NB_MODE_PLAIN : s->p.mode; }

// The coarse and learned traversals take their 64-body groups from d.gorder,
// most expensive first. The order for a pass was worked out on the order
// stream after the previous pass (this stream waits for it; the first pass
// works it out here), and the order for the next pass follows this one there:
// it reads the costs this pass writes, and nothing else touches d.gcost or
// d.gorder in between.
int launch_coarse_ordered(nb_sim *s, int out, int fmode)
{
    const nb::Dev &d = s->d;
    if (!s->ord) {
        HIP_TRY(s, hipStreamCreateWithFlags(&s->ord, hipStreamNonBlocking));
        HIP_TRY(s, hipEventCreateWithFlags(&s->ord_in, hipEventDisableTiming));
        HIP_TRY(s, hipEventCreateWithFlags(&s->ord_out, hipEventDisableTiming));
    }
    if (s->ord_ready)
        HIP_TRY(s, hipStreamWaitEvent(s->stream, s->ord_out, 0));
    else
        HIP_TRY(s, nb::launch_group_order(d, s->stream));
    HIP_TRY(s, nb::launch_coarse(d, s->L, out, nullptr, d.gorder, (d.n + 63) / 64, nullptr, nullptr, nullptr, fmode,
                                 s->stream));
    HIP_TRY(s, hipEventRecord(s->ord_in, s->stream));
    HIP_TRY(s, hipStreamWaitEvent(s->ord, s->ord_in, 0));
    HIP_TRY(s, nb::launch_group_order(d, s->ord));
    HIP_TRY(s, hipEventRecord(s->ord_out, s->ord));
    s->ord_ready = true;
    return NB_OK;
}

// Force stage with an explicit method; fmode as nb::launch_force (0 store,
// 1 fused update, 2 probe).
int launch_force_method(nb_sim *s, int method, int fmode)
{
    const nb::Dev &d = s->d;
    hipError_t e = hipSuccess;
    int r = NB_OK;
    switch (method) {
    case NB_MODE_PLAIN: e = nb::launch_force(d, s->p.vote_width, fmode, s->stream); break;
    case NB_MODE_COARSE: r = launch_coarse_ordered(s, nb::kOutFinish, fmode); break;
    case NB_MODE_LEARNED:
        if (!s->have_model) return fail(s, NB_E_MODEL, "learned mode without a model (nb_train or nb_set_model)");
        r = launch_coarse_ordered(s, nb::kOutLearned, fmode);
        break;
    default: return fail(s, NB_E_ARG, "unknown force method %d", method);
    }
    if (e != hipSuccess) return fail(s, NB_E_HIP, "force launch: %s", hipGetErrorString(e));
    // the fused update writes every body's new record at its sorted place too
    if (!r && fmode == 1) s->staged = true;
    return r;
}

int launch_stage(nb_sim *s, int stage)
{
    const nb::Dev &d = s->d;
    hipError_t e = hipSuccess;
    const int fused = (s->p.flags & NB_FUSED_UPDATE) ? 1 : 0;
    switch (stage) {
    case NB_STAGE_BBOX: e = nb::launch_bbox(d, s->bbox_blocks, s->stream); break;
    case NB_STAGE_BUILD:
        e = nb::launch_build_keys(d, s->keys_warm, s->keys_warm && s->staged, s->stream);
        s->sort_pending = e == hipSuccess;
        break;
    case NB_STAGE_SUMMARIZE: {
        int r = prepare_levels(s);
        if (r) return r;
        e = nb::launch_summarize_block(d, s->split, nb::kMaxDepth + 1, s->stream);
        for (int L = s->split - 1; L >= s->head && e == hipSuccess; --L)
            e = nb::launch_summarize_level(d, L, hint_cells(s, L), s->stream);
        if (e == hipSuccess && s->head > 0) e = nb::launch_summarize_block(d, 0, s->head, s->stream);
        break;
    }
    case NB_STAGE_SORT:
        // the build's sorted body order is the octant-path order (sort.cl:12-76):
        // its buffer becomes sorted[], and the last order's buffer takes its
        // place for the next build (launches already queued keep the pointers
        // they were given)
        if (s->sort_pending) {
            std::swap(s->d.sorted, s->d.kb_val[0]);
            s->sort_pending = false;
        }
        s->keys_warm = true; // sorted[] holds this tree's body order for the next build
        s->staged = false;   // until the fused update writes bnode[] in that order
        break;
    case NB_STAGE_FORCE: return launch_force_method(s, force_method(s), fused);
    case NB_STAGE_INTEGRATE:
        if (!fused) {
            e = nb::launch_integrate(d, s->stream);
            s->staged = false;
        }
        break;
    default: return fail(s, NB_E_ARG, "stage %d out of range 1..6", stage);
    }
    if (e != hipSuccess) return fail(s, NB_E_HIP, "stage %d launch: %s", stage, hipGetErrorString(e));
    return NB_OK;
}

// Moment buffers of the quadrupole summarize (learned and coarse modes, training).
int ensure_quad(nb_sim *s)
{
    nb::Dev &d = s->d;
    const size_t cap = (size_t)d.cap;
    if (!d.cq64) HIP_TRY(s, dalloc(&d.cq64, 8 * cap));
    if (!d.cqa) HIP_TRY(s, dalloc(&d.cqa, 8 * cap));
    if (!d.cqb) HIP_TRY(s, dalloc(&d.cqb, 8 * cap));
    if (!d.cmac) HIP_TRY(s, dalloc(&d.cmac, 8 * cap));
    return NB_OK;
}

int ensure_learn(nb_sim *s)
{
    if (!s->mlpbuf) HIP_TRY(s, hipMalloc(&s->mlpbuf, nb::kMlpBufBytes));
    return ensure_quad(s);
}

// With the handle's device current: the device address of `scratch` there,
// looked up on the first call for that device.
int init_scratch(nb_sim *s)
{
    if (s->device < 0 || s->device >= kMaxDevices)
        return fail(s, NB_E_ARG, "device %d is beyond the %d devices of the staging table", s->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[s->device]) return NB_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(s, NB_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[s->device] = static_cast<float4 *>(p);
    return NB_OK;
}

float4 *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Waits for everything queued on the handle's streams (the main stream, the
// sampled check's side stream, the launch-order stream); errors are ignored:
// this runs before host data overwrites device buffers that queued kernels may
// still read, and on the error returns of nb_step.
void drain(nb_sim *s)
{
    for (hipStream_t st : {s->stream, s->side, s->ord})
        if (st) (void)hipStreamSynchronize(st);
}

// Host -> device copy ordered on the handle's stream, complete on return.
int upload(nb_sim *s, void *dst, const void *src, size_t bytes)
{
    HIP_TRY(s, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}

// Learned-mode control state back to what a new handle starts with: the
// configured alpha, no fallback, the check schedule and its sample seed at
// their start.
void reset_control(nb_sim *s)
{
    s->alpha_now = s->p.coarse_alpha;
    s->d.alpha2_coarse = s->p.coarse_alpha > 0.0f ? (float)((double)s->p.coarse_alpha * (double)s->p.coarse_alpha) : 0.0f;
    s->fell_back = false;
    s->fail_streak = 0;
    s->steps_since_model = 0;
    s->check_seed = 0x5eed;
}

uint16_t bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Validates a model blob, packs the inference network and uploads it.
int install_model(nb_sim *s, const float *blob)
{
    for (int i = 0; i < NB_MODEL_FLOATS; ++i)
        if (!std::isfinite(blob[i])) return fail(s, NB_E_MODEL, "model value %d is not finite", i);
    const float *P = blob;
    const float *xmu = blob + nb::kMlpParams, *xsd = xmu + nb::kFeat;
    const float *ymu = xsd + nb::kFeat, *ysd = ymu + nb::kOut;
    for (int i = 0; i < nb::kFeat; ++i)
        if (!(xsd[i] > 0.0f)) return fail(s, NB_E_MODEL, "feature scale %d is not positive", i);
    for (int i = 0; i < nb::kOut; ++i)
        if (!(ysd[i] > 0.0f)) return fail(s, NB_E_MODEL, "target scale %d is not positive", i);
    int r = ensure_learn(s);
    if (r) return r;
    std::vector<unsigned char> buf(nb::kMlpBufBytes, 0);
    uint16_t *w1f = reinterpret_cast<uint16_t *>(buf.data());
    uint16_t *w2f = w1f + nb::kW1Frag;
    uint16_t *w3f = w2f + nb::kW2Frag;
    float *fb = reinterpret_cast<float *>(w3f + nb::kW3Frag);
    float *b1 = fb, *b2 = fb + 64, *b3 = fb + 128, *mu = fb + 144, *rs = fb + 168;
    // fragment order of the transposed network (nb_internal.h, struct Mlp):
    // lane l holds row l % 16 of its 16-row block and k group g = l / 16
    auto pk = [](int kb, int g, int i) { return 32 * kb + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4)); };
    for (int rb = 0; rb < 4; ++rb)
        for (int l = 0; l < 64; ++l)
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * (l / 16) + i, nn = 16 * rb + l % 16;
                w1f[(rb * 64 + l) * 8 + i] = k < nb::kFeat ? bf16_rne(P[k * 64 + nn]) : 0;
            }
    for (int rb = 0; rb < 4; ++rb)
        for (int kb = 0; kb < 2; ++kb)
            for (int l = 0; l < 64; ++l)
                for (int i = 0; i < 8; ++i) {
                    const int k = pk(kb, l / 16, i), nn = 16 * rb + l % 16;
                    w2f[((rb * 2 + kb) * 64 + l) * 8 + i] = bf16_rne(P[nb::kOffW2 + k * 64 + nn]);
                }
    for (int kb = 0; kb < 2; ++kb)
        for (int l = 0; l < 64; ++l)
            for (int i = 0; i < 8; ++i) {
                const int k = pk(kb, l / 16, i), o = l % 16;
                w3f[(kb * 64 + l) * 8 + i] = o < nb::kOut ? bf16_rne(P[nb::kOffW3 + k * 3 + o] * ysd[o]) : 0;
            }
    for (int i = 0; i < 64; ++i) {
        b1[i] = P[nb::kOffB1 + i];
        b2[i] = P[nb::kOffB2 + i];
    }
    for (int i = 0; i < nb::kOut; ++i) b3[i] = P[nb::kOffB3 + i] * ysd[i] + ymu[i];
    for (int i = 0; i < nb::kFeat; ++i) {
        mu[i] = xmu[i];
        rs[i] = 1.0f / xsd[i];
    }
    drain(s);
    r = upload(s, s->mlpbuf, buf.data(), buf.size());
    if (r) return r;
    const unsigned char *base = static_cast<const unsigned char *>(s->mlpbuf);
    const size_t fo = (size_t)(reinterpret_cast<unsigned char *>(fb) - buf.data());
    s->L.mlp.w1f = reinterpret_cast<const uint16_t *>(base);
    s->L.mlp.w2f = s->L.mlp.w1f + nb::kW1Frag;
    s->L.mlp.w3f = s->L.mlp.w2f + nb::kW2Frag;
    const float *dfb = reinterpret_cast<const float *>(base + fo);
    s->L.mlp.b1 = dfb;
    s->L.mlp.b2 = dfb + 64;
    s->L.mlp.b3 = dfb + 128;
    s->L.mlp.xmu = dfb + 144;
    s->L.mlp.xrs = dfb + 168;
    s->model.assign(blob, blob + NB_MODEL_FLOATS);
    s->have_model = true;
    reset_control(s);
    return NB_OK;
}

void free_all(nb_sim *s)
{
    nb::Dev &d = s->d;
    if (s->ord) (void)hipStreamSynchronize(s->ord); // it may still be writing d.gorder
    void *ptrs[] = {d.ctl,   d.pm4,   d.vel4,   d.acc4,          d.pos64,    d.slot,
                    d.cinfo, d.count, d.com64,         d.cchild,   d.lvl_list, d.clvl,
                    d.lvl_blockhist,  d.cdata,  d.bnode,  d.sorted, d.probe, d.bbox_part, d.gcost, d.gorder,
                    s->chk_idx, s->chk_part, s->chk_out, d.cq64, d.cqa, d.cqb, d.cmac,
                    s->mlpbuf, s->chk_acc, s->vtx, d.kb_key[0], d.kb_key[1], d.kb_val[0], d.kb_val[1],
                    d.kb_deep, d.kb_hist, d.kb_offs, d.kb_tot, d.kb_ipar, d.kb_lpar, d.kb_lo,
                    d.kb_own, d.kb_ktop, d.kb_obase, d.kb_scan, d.kb_tdep, d.kb_spl, d.kb_bcnt, d.kb_bfill, d.kb_bstart, d.kb_big};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : s->tev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : s->chk_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->chk_go) (void)hipEventDestroy(s->chk_go);
    if (s->chk_host) (void)hipHostFree(s->chk_host);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->ord_in) (void)hipEventDestroy(s->ord_in);
    if (s->ord_out) (void)hipEventDestroy(s->ord_out);
    if (s->ord) (void)hipStreamDestroy(s->ord);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

template <class T> int fetch(nb_sim *s, std::vector<T> &out, const T *src, size_t count)
{
    out.resize(count);
    HIP_TRY(s, hipMemcpyAsync(out.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}

int ensure_check_buffers(nb_sim *s, int ns)
{
    if (s->chk_cap >= ns) return NB_OK;
    if (s->chk_idx) (void)hipFree(s->chk_idx);
    if (s->chk_part) (void)hipFree(s->chk_part);
    if (s->chk_out) (void)hipFree(s->chk_out);
    s->chk_idx = nullptr;
    s->chk_part = s->chk_out = nullptr;
    s->chk_cap = 0;
    HIP_TRY(s, hipMalloc((void **)&s->chk_idx, (size_t)ns * sizeof(int)));
    HIP_TRY(s, hipMalloc((void **)&s->chk_part, (size_t)ns * 64 * sizeof(double4)));
    HIP_TRY(s, hipMalloc((void **)&s->chk_out, (size_t)ns * sizeof(double4)));
    if (s->chk_acc) (void)hipFree(s->chk_acc);
    s->chk_acc = nullptr;
    HIP_TRY(s, hipMalloc((void **)&s->chk_acc, (size_t)ns * sizeof(float4)));
    s->chk_cap = ns;
    return NB_OK;
}

// Device allocations that live for one API call; freed after the stream drains.
struct Scratch {
    nb_sim *s;
    std::vector<void *> ptrs;
    explicit Scratch(nb_sim *sim) : s(sim) {}
    ~Scratch()
    {
        (void)hipStreamSynchronize(s->stream);
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T> int get(T **p, size_t count)
    {
        *p = nullptr;
        hipError_t e = hipMalloc((void **)p, count * sizeof(T) + 64);
        if (e != hipSuccess) return fail(s, NB_E_HIP, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
        ptrs.push_back(*p);
        return NB_OK;
    }
};

// Training buffers for S samples (X, T given or allocated) and batch B.
int train_buffers(Scratch &m, nb::Train &t, int S, int B, int iters)
{
    int r = NB_OK;
    for (float **f : {&t.P, &t.G, &t.M, &t.V})
        if (!r) r = m.get(f, (size_t)nb::kMlpParams);
    for (float **f : {&t.Z1, &t.H1, &t.Z2, &t.H2, &t.dZ1, &t.dZ2})
        if (!r) r = m.get(f, (size_t)B * nb::kHid);
    if (!r) r = m.get(&t.dY, (size_t)B * nb::kOut);
    if (!r) r = m.get(&t.gpart, (size_t)nb::kGradChunks * (nb::kHid + 1) * nb::kHid);
    if (!r) r = m.get(&t.loss_part, (size_t)(B + 255) / 256);
    if (!r) r = m.get(&t.loss_hist, (size_t)std::max(iters, 1));
    t.S = S;
    return r;
}

uint64_t splitmix64(uint64_t &x)
{
    x += 0x9e3779b97f4a7c15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Starts a check step's FP64 direct sum on the side stream. It reads only the
// positions at the start of the step, which nothing changes before the
// integrate, so it runs beside the tree stages and the traversal; the step
// joins it (an event) before it compares.
int start_check(nb_sim *s)
{
    const int n = s->d.n;
    const int ns = std::min(s->p.check_samples, n);
    int r = ensure_check_buffers(s, ns);
    if (r) return r;
    if (!s->side) HIP_TRY(s, hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    for (auto &ev : s->chk_ev)
        if (!ev) HIP_TRY(s, hipEventCreate(&ev));
    if (!s->chk_go) HIP_TRY(s, hipEventCreateWithFlags(&s->chk_go, hipEventDisableTiming));
    if (s->chk_host_cap < ns) {
        // no upload from it is in flight: the last check step waited for its own
        if (s->chk_host) (void)hipHostFree(s->chk_host);
        s->chk_host = nullptr;
        s->chk_host_cap = 0;
        HIP_TRY(s, hipHostMalloc((void **)&s->chk_host, (size_t)ns * sizeof(int32_t)));
        s->chk_host_cap = ns;
    }
    for (int i = 0; i < ns; ++i) s->chk_host[i] = (int32_t)(splitmix64(s->check_seed) % (uint64_t)n);
    HIP_TRY(s, hipEventRecord(s->chk_go, s->stream));
    HIP_TRY(s, hipStreamWaitEvent(s->side, s->chk_go, 0));
    HIP_TRY(s, hipMemcpyAsync(s->chk_idx, s->chk_host, (size_t)ns * sizeof(int32_t), hipMemcpyHostToDevice, s->side));
    HIP_TRY(s, hipEventRecord(s->chk_ev[0], s->side));
    hipError_t e = nb::launch_direct64(s->d, s->chk_idx, ns, s->chk_part, s->chk_out, false, s->side);
    if (e == hipSuccess) e = hipEventRecord(s->chk_ev[1], s->side);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "sampled check launch: %s", hipGetErrorString(e));
    s->chk_ns = ns;
    return NB_OK;
}

// Force stage of a check step (its direct sum started with the step): the
// learned forces are stored (velocity correction included), the check_samples
// bodies (drawn with replacement) are compared with the FP64 direct sum, and
// if the median relative error is above check_tol the step's forces are
// recomputed with plain theta. The second pass corrects the velocities by
// (a_plain - a_learned) dt / 2, so the step ends exactly as a plain step
// would. The fallback holds until a new model is installed.
int checked_force(nb_sim *s)
{
    int r = launch_force_method(s, NB_MODE_LEARNED, 0);
    if (r) return r;
    const int ns = s->chk_ns;
    HIP_TRY(s, hipStreamWaitEvent(s->stream, s->chk_ev[1], 0));
    hipError_t e = nb::launch_gather_acc(s->d, s->chk_idx, ns, s->chk_acc, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "sampled check launch: %s", hipGetErrorString(e));
    std::vector<double4> ref;
    std::vector<float4> got;
    r = fetch(s, ref, (const double4 *)s->chk_out, (size_t)ns);
    if (!r) r = fetch(s, got, (const float4 *)s->chk_acc, (size_t)ns);
    if (r) return r;
    float ms = 0.f; // both events have completed: the stream waited for the second
    HIP_TRY(s, hipEventElapsedTime(&ms, s->chk_ev[0], s->chk_ev[1]));
    s->check_ms += ms;
    ++s->checks;
    {
        // a device error earlier in the step left the accelerations unset:
        // report it instead of judging them
        nb::Ctl c;
        r = read_ctl(s, &c);
        if (r) return r;
        update_hint(s, c);
        if (c.error) return check_device_error(s);
    }
    std::vector<double> err((size_t)ns);
    for (int i = 0; i < ns; ++i) {
        const double dx = got[i].x - ref[i].x, dy = got[i].y - ref[i].y, dz = got[i].z - ref[i].z;
        const double den = std::sqrt(ref[i].x * ref[i].x + ref[i].y * ref[i].y + ref[i].z * ref[i].z);
        err[i] = std::sqrt(dx * dx + dy * dy + dz * dz) / std::max(den, 1e-300);
    }
    std::nth_element(err.begin(), err.begin() + (ns - 1) / 2, err.end());
    s->last_check_median = err[(size_t)(ns - 1) / 2];
    if (s->last_check_median > (double)s->p.check_tol) {
        r = launch_force_method(s, NB_MODE_PLAIN, 0);
        if (r) return r;
        ++s->fallbacks;
        if (s->p.check_adapt && ++s->fail_streak < 3 && s->alpha_now > 0.f) {
            // The coarse error grows about as alpha^0.45 (the benchmark
            // workload: 0.0175 -> 0.025 raises the median by 17 %): aim at
            // 85 % of the tolerance, at most halving alpha per failed check.
            double shrink = std::pow(0.85 * (double)s->p.check_tol / s->last_check_median, 1.0 / 0.45);
            shrink = std::min(0.95, std::max(0.5, shrink));
            s->alpha_now = (float)((double)s->alpha_now * shrink);
            s->d.alpha2_coarse = (float)((double)s->alpha_now * (double)s->alpha_now);
        } else {
            s->fell_back = true;
        }
    } else {
        s->fail_streak = 0;
    }
    return NB_OK;
}

} // namespace

extern "C" {

void nb_default_params(nb_params *p, int32_t n_bodies)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->n_bodies = n_bodies;
    p->theta = std::sqrt(1.5f); // THETA 1.5 = theta^2 (calculateforce.cl:16)
    p->softening = 0.05f;       // EPSILON 0.0025 = softening^2 (calculateforce.cl:14)
    p->dt = 0.025f;             // TIMESTEP (calculateforce.cl:18)
    p->vote_width = 16;         // WARPSIZE (GPUBarnesHutNBodySimulation.java:47)
    p->flags = 0;
    p->device = 0;
    p->pool_factor = 2.0f;      // NUMBER_OF_NODES = 2 * NBODIES (java:220)
    p->mode = NB_MODE_PLAIN;
    p->coarse_theta = 0.95f;
    p->coarse_alpha = 0.02f;
    p->check_tol = 1e-3f;
    p->check_every = 0;
    p->check_samples = 4096;
}

int nb_create(const nb_params *p, nb_sim **out)
{
    if (!out) return NB_E_ARG;
    *out = nullptr;
    int v = validate(p);
    if (v) {
        g_create_error = "invalid nb_params";
        return v;
    }
    nb_sim *s = new nb_sim();
    s->p = *p;
    const int n = p->n_bodies;
    auto bail = [&](int code) {
        g_create_error = s->err.empty() ? nb_status_string(code) : s->err;
        free_all(s);
        delete s;
        return code;
    };
    hipError_t e = hipSetDevice(p->device);
    if (e != hipSuccess) {
        fail(s, NB_E_HIP, "hipSetDevice(%d): %s", p->device, hipGetErrorString(e));
        return bail(NB_E_HIP);
    }
    s->device = p->device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) return bail(NB_E_HIP);
    s->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bail(NB_E_HIP);

    nb::Dev &d = s->d;
    d.n = n;
    d.cap = (int)pool_cells(p);
    d.theta2 = p->theta * p->theta;       // THETA = theta^2
    d.eps2 = p->softening * p->softening; // EPSILON = softening^2
    d.dt = p->dt;
    {
        // the oracle derives the same two FP32 constants in the same way
        const float theta2c = p->coarse_theta * p->coarse_theta;
        d.it2_coarse = 1.0f / theta2c;
        d.alpha2_coarse = p->coarse_alpha > 0.0f ? (float)((double)p->coarse_alpha * (double)p->coarse_alpha) : 0.0f;
        s->alpha_now = p->coarse_alpha;
    }
    d.fp64 = (p->flags & NB_FP64_POSITIONS) ? 1 : 0;
    d.quad = p->mode != NB_MODE_PLAIN ? 1 : 0;
    const size_t cap = (size_t)d.cap;
    bool ok = dalloc(&d.ctl, 1) == hipSuccess;
    for (float4 **f : {&d.pm4, &d.vel4, &d.acc4}) ok = ok && dalloc(f, (size_t)n) == hipSuccess;
    if (d.fp64) ok = ok && dalloc(&d.pos64, (size_t)n) == hipSuccess;
    ok = ok && dalloc(&d.slot, 8 * cap) == hipSuccess;
    ok = ok && dalloc(&d.cinfo, cap) == hipSuccess;
    ok = ok && dalloc(&d.count, cap) == hipSuccess;
    ok = ok && dalloc(&d.com64, cap) == hipSuccess;
    ok = ok && dalloc(&d.cchild, 8 * cap) == hipSuccess;
    ok = ok && dalloc(&d.lvl_list, cap) == hipSuccess && dalloc(&d.clvl, cap) == hipSuccess;
    ok = ok && dalloc(&d.lvl_blockhist, (size_t)nb::kLevelBlocks * nb::kMaxDepth) == hipSuccess;
    ok = ok && dalloc(&d.cdata, 8 * cap) == hipSuccess;
    ok = ok && dalloc(&d.bnode, (size_t)n) == hipSuccess && dalloc(&d.probe, (size_t)n) == hipSuccess;
    ok = ok && dalloc(&d.sorted, (size_t)n) == hipSuccess;
    ok = ok && dalloc(&d.bbox_part, (size_t)6 * nb::kBBoxBlocks) == hipSuccess;
    const size_t ngroups = ((size_t)n + 63) / 64;
    ok = ok && dalloc(&d.gcost, ngroups) == hipSuccess && dalloc(&d.gorder, ngroups) == hipSuccess;
    {
        const size_t ktiles = ((size_t)n + nb::kKeyTile - 1) / nb::kKeyTile;
        const size_t stiles = ((size_t)n + nb::kKeyScanTile - 1) / nb::kKeyScanTile;
        for (int b = 0; b < 2; ++b)
            ok = ok && dalloc(&d.kb_key[b], (size_t)n) == hipSuccess && dalloc(&d.kb_val[b], (size_t)n) == hipSuccess;
        ok = ok && dalloc(&d.kb_deep, 3 * (size_t)n) == hipSuccess;
        ok = ok && dalloc(&d.kb_hist, 256 * ktiles) == hipSuccess && dalloc(&d.kb_offs, 256 * ktiles) == hipSuccess;
        ok = ok && dalloc(&d.kb_tot, 256) == hipSuccess;
        for (int **a : {&d.kb_ipar, &d.kb_lpar, &d.kb_lo, &d.kb_own, &d.kb_ktop, &d.kb_obase})
            ok = ok && dalloc(a, (size_t)n) == hipSuccess;
        ok = ok && dalloc(&d.kb_scan, stiles) == hipSuccess && dalloc(&d.kb_tdep, stiles) == hipSuccess;
        const size_t S = (size_t)nb::bucket_samples(n);
        ok = ok && dalloc(&d.kb_spl, S) == hipSuccess && dalloc(&d.kb_bcnt, S + 1) == hipSuccess &&
             dalloc(&d.kb_bfill, S + 1) == hipSuccess && dalloc(&d.kb_bstart, S + 2) == hipSuccess &&
             dalloc(&d.kb_big, S + 2) == hipSuccess;
    }
    if (!ok) {
        fail(s, NB_E_HIP, "device allocation failed for %d bodies / %d cells", n, d.cap);
        return bail(NB_E_HIP);
    }
    if (d.quad && ensure_learn(s) != NB_OK) return bail(NB_E_HIP);
    if (const int rs = init_scratch(s)) return bail(rs);
    // epoch 0 tags every slot as stale; the first bounding box moves to epoch 1
    ok = hipMemsetAsync(d.slot, 0, 8 * cap * sizeof(unsigned long long), s->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(d.cinfo, 0xff, cap * sizeof(int4), s->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(d.count, 0, cap * sizeof(int), s->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(d.gcost, 0, ngroups * sizeof(int), s->stream) == hipSuccess;
    nb::Ctl c = initial_ctl(s, 0);
    ok = ok && hipMemcpyAsync(d.ctl, &c, sizeof c, hipMemcpyHostToDevice, s->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(s->stream) == hipSuccess;
    if (!ok) return bail(NB_E_HIP);

    s->bbox_blocks = std::max(1, std::min(nb::kBBoxBlocks, (n + 2047) / 2048));
    *out = s;
    return NB_OK;
}

void nb_destroy(nb_sim *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

int nb_set_state(nb_sim *s, const float *x, const float *y, const float *z, const float *vx,
                 const float *vy, const float *vz, const float *m)
{
    if (!s) return NB_E_ARG;
    if (!x || !y || !z || !vx || !vy || !vz || !m) return fail(s, NB_E_ARG, "null state array");
    HIP_TRY(s, hipSetDevice(s->device));
    // nothing queued earlier may still read the buffers overwritten below
    drain(s);
    nb::Dev &d = s->d;
    const size_t n = (size_t)d.n;
    int r = NB_OK;
    {
        std::vector<float4> pm(n), vel(n), zero(n, make_float4(0.f, 0.f, 0.f, 0.f));
        for (size_t i = 0; i < n; ++i) {
            pm[i] = make_float4(x[i], y[i], z[i], m[i]);
            vel[i] = make_float4(vx[i], vy[i], vz[i], 0.f);
        }
        r = upload(s, d.pm4, pm.data(), n * sizeof(float4));
        if (!r) r = upload(s, d.vel4, vel.data(), n * sizeof(float4));
        if (!r) r = upload(s, d.acc4, zero.data(), n * sizeof(float4));
        if (r) return r;
    }
    if (d.fp64) {
        std::vector<double4> p64(n);
        for (size_t i = 0; i < n; ++i) p64[i] = make_double4(x[i], y[i], z[i], 0.0);
        r = upload(s, d.pos64, p64.data(), n * sizeof(double4));
        if (r) return r;
    }
    // sorted[] before the first tree: the identity
    {
        std::vector<int> ident((size_t)d.n);
        for (int i = 0; i < d.n; ++i) ident[i] = i;
        r = upload(s, d.sorted, ident.data(), (size_t)d.n * sizeof(int));
        if (r) return r;
    }
    // keep the epoch running so that no slot of an earlier tree can look current
    nb::Ctl cur;
    r = read_ctl(s, &cur);
    if (r) return r;
    nb::Ctl c = initial_ctl(s, cur.epoch);
    HIP_TRY(s, hipMemcpyAsync(d.ctl, &c, sizeof c, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    s->inter_before = 0;
    s->cell_inter_before = 0;
    s->hint_valid = false; // the first step reads its level sizes back
    s->ord_ready = false;  // the launch order of the coarse traversal belongs to the old bodies
    s->keys_warm = false;  // and so does sorted[]: the next build sorts from scratch
    s->staged = false;
    s->sort_pending = false;
    reset_control(s);
    s->have_state = true;
    s->err.clear();
    return NB_OK;
}

int nb_run_stage(nb_sim *s, int32_t stage)
{
    if (!s) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    HIP_TRY(s, hipSetDevice(s->device));
    int r = launch_stage(s, stage);
    if (r) return r;
    return check_device_error(s);
}

static int step_loop(nb_sim *s, int32_t nsteps);

int nb_step(nb_sim *s, int32_t nsteps)
{
    if (!s) return NB_E_ARG;
    if (nsteps < 0) return fail(s, NB_E_ARG, "negative step count");
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    HIP_TRY(s, hipSetDevice(s->device));
    if (force_method(s) == NB_MODE_LEARNED && !s->have_model)
        return fail(s, NB_E_MODEL, "learned mode without a model (nb_train or nb_set_model)");
    const int r = step_loop(s, nsteps);
    // a failed step may have left launches queued on any of the streams: none
    // may still run when the call returns
    if (r) drain(s);
    return r;
}

static int step_loop(nb_sim *s, int32_t nsteps)
{
    // with timing on, every stage of every step is bracketed by events on the
    // handle's stream; stage_ms reports the per-step average
    const size_t nev = s->timing ? (size_t)7 * nsteps : 0;
    while (s->tev.size() < nev) {
        hipEvent_t e;
        HIP_TRY(s, hipEventCreate(&e));
        s->tev.push_back(e);
    }
    const bool learned = s->p.mode == NB_MODE_LEARNED;
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t *ev = s->timing ? s->tev.data() + 7 * (size_t)k : nullptr;
        // the sampled FP64 check of the learned mode runs on every
        // check_every-th step: force without the fused update, check, fall
        // back to plain theta if the median error is above check_tol
        const bool check = learned && !s->fell_back && s->p.check_every > 0 && s->p.check_tol > 0.0f &&
                           s->p.check_samples > 0 && (s->steps_since_model % s->p.check_every) == 0;
        if (check) {
            const int r = start_check(s);
            if (r) return r;
        }
        for (int st = NB_STAGE_BBOX; st <= NB_STAGE_INTEGRATE; ++st) {
            if (ev) HIP_TRY(s, hipEventRecord(ev[st - 1], s->stream));
            int r;
            if (check && st == NB_STAGE_FORCE)
                r = checked_force(s);
            else if (check && st == NB_STAGE_INTEGRATE) {
                r = nb::launch_integrate(s->d, s->stream) == hipSuccess ? NB_OK
                                                                         : fail(s, NB_E_HIP, "integrate launch");
                s->staged = false;
            }
            else
                r = launch_stage(s, st);
            if (r) return r;
        }
        if (ev) HIP_TRY(s, hipEventRecord(ev[6], s->stream));
        if (learned) ++s->steps_since_model;
    }
    int r = check_device_error(s);
    if (s->timing && nsteps > 0) {
        for (int st = 0; st < 6; ++st) {
            double tot = 0.0;
            for (int k = 0; k < nsteps; ++k) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, s->tev[7 * (size_t)k + st], s->tev[7 * (size_t)k + st + 1]) == hipSuccess)
                    tot += ms;
            }
            s->stage_ms[st] = tot / nsteps;
        }
    }
    return r;
}

int nb_copy_vertices(nb_sim *s, float *pos4, float *vel4)
{
    if (!s) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    if (!pos4 && !vel4) return NB_OK;
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->d.n, bytes = n * sizeof(float4);
    // the staging buffer is taken once and reused by every later call (the
    // renderer calls this after every step)
    if (!s->vtx) HIP_TRY(s, dalloc(&s->vtx, 2 * n));
    float4 *dp = pos4 ? s->vtx : nullptr, *dv = vel4 ? s->vtx + n : nullptr;
    hipError_t e = nb::launch_copy_vertices(s->d, 0, (int)n, dp, dv, s->stream);
    if (e == hipSuccess && dp) e = hipMemcpyAsync(pos4, dp, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && dv) e = hipMemcpyAsync(vel4, dv, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "copy_vertices: %s", hipGetErrorString(e));
    return NB_OK;
}

int nb_copy_vertices_noalloc(nb_sim *s, float *pos4, float *vel4)
{
    if (!s) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    if (!pos4 && !vel4) return NB_OK;
    HIP_TRY(s, hipSetDevice(s->device));
    float4 *stage = scratch_ptr(s->device);
    if (!stage) return fail(s, NB_E_STATE, "no staging address for device %d", s->device);
    const int n = s->d.n;
    // one chunk at a time through the staging array; the next chunk's kernel
    // is ordered after the previous chunk's copies on the same stream
    for (int first = 0; first < n; first += kNoAllocChunk) {
        const int cnt = std::min(kNoAllocChunk, n - first);
        const size_t bytes = (size_t)cnt * sizeof(float4);
        float4 *dp = pos4 ? stage : nullptr, *dv = vel4 ? stage + kNoAllocChunk : nullptr;
        hipError_t e = nb::launch_copy_vertices(s->d, first, cnt, dp, dv, s->stream);
        if (e == hipSuccess && dp) e = hipMemcpyAsync(pos4 + 4 * (size_t)first, dp, bytes, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess && dv) e = hipMemcpyAsync(vel4 + 4 * (size_t)first, dv, bytes, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(s->stream);
            return fail(s, NB_E_HIP, "copy_vertices_noalloc: %s", hipGetErrorString(e));
        }
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}

int32_t nb_num_bodies(const nb_sim *s) { return s ? s->d.n : 0; }

int nb_get_state(nb_sim *s, float *x, float *y, float *z, float *vx, float *vy, float *vz)
{
    if (!s) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->d.n;
    std::vector<float4> p, v;
    int r = fetch(s, p, s->d.pm4, n);
    if (!r) r = fetch(s, v, s->d.vel4, n);
    if (r) return r;
    for (size_t i = 0; i < n; ++i) {
        if (x) x[i] = p[i].x;
        if (y) y[i] = p[i].y;
        if (z) z[i] = p[i].z;
        if (vx) vx[i] = v[i].x;
        if (vy) vy[i] = v[i].y;
        if (vz) vz[i] = v[i].z;
    }
    return NB_OK;
}

int nb_get_accel(nb_sim *s, float *ax, float *ay, float *az)
{
    if (!s) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->d.n;
    std::vector<float4> a;
    int r = fetch(s, a, s->d.acc4, n);
    if (r) return r;
    for (size_t i = 0; i < n; ++i) {
        if (ax) ax[i] = a[i].x;
        if (ay) ay[i] = a[i].y;
        if (az) az[i] = a[i].z;
    }
    return NB_OK;
}

int nb_get_positions64(nb_sim *s, double *x, double *y, double *z)
{
    if (!s) return NB_E_ARG;
    if (!s->d.fp64) return fail(s, NB_E_STATE, "NB_FP64_POSITIONS is not set");
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->d.n;
    std::vector<double4> p;
    int r = fetch(s, p, s->d.pos64, n);
    if (r) return r;
    for (size_t i = 0; i < n; ++i) {
        if (x) x[i] = p[i].x;
        if (y) y[i] = p[i].y;
        if (z) z[i] = p[i].z;
    }
    return NB_OK;
}

int nb_get_root(nb_sim *s, float *out4)
{
    if (!s || !out4) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    nb::Ctl c;
    int r = read_ctl(s, &c);
    if (r) return r;
    out4[0] = c.root[0];
    out4[1] = c.root[1];
    out4[2] = c.root[2];
    out4[3] = c.radius;
    return NB_OK;
}

int nb_get_sorted(nb_sim *s, int32_t *sorted)
{
    if (!s || !sorted) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    HIP_TRY(s, hipMemcpyAsync(sorted, s->d.sorted, (size_t)s->d.n * sizeof(int), hipMemcpyDeviceToHost,
                              s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}

int nb_get_paths(nb_sim *s, uint64_t *paths3, int32_t *depth)
{
    if (!s || !paths3 || !depth) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->d.n;
    unsigned long long *dp = nullptr;
    int *dd = nullptr;
    HIP_TRY(s, hipMalloc((void **)&dp, 3 * n * sizeof(unsigned long long)));
    hipError_t e = hipMalloc((void **)&dd, n * sizeof(int));
    if (e == hipSuccess) e = nb::launch_paths(s->d, dp, dd, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(paths3, dp, 3 * n * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(depth, dd, n * sizeof(int), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dp);
    if (dd) (void)hipFree(dd);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "get_paths: %s", hipGetErrorString(e));
    return NB_OK;
}

int nb_get_cells(nb_sim *s, int32_t cap, int32_t *ncells, int32_t *count, float *x, float *y, float *z,
                 float *m)
{
    if (!s || !ncells) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    nb::Ctl c;
    int r = read_ctl(s, &c);
    if (r) return r;
    const int live = c.live_cells;
    *ncells = live;
    if (cap < live) return fail(s, NB_E_ARG, "cell buffers hold %d of %d cells", cap, live);
    std::vector<int> list, cnt;
    std::vector<double4> com;
    r = fetch(s, list, s->d.lvl_list, (size_t)live);
    if (!r) r = fetch(s, cnt, s->d.count, (size_t)s->d.cap);
    if (!r) r = fetch(s, com, s->d.com64, (size_t)s->d.cap);
    if (r) return r;
    for (int i = 0; i < live; ++i) {
        const int k = list[i];
        if (count) count[i] = cnt[k];
        if (x) x[i] = (float)com[k].x;
        if (y) y[i] = (float)com[k].y;
        if (z) z[i] = (float)com[k].z;
        if (m) m[i] = (float)com[k].w;
    }
    return NB_OK;
}

// ---- force-error measurement ------------------------------------------------

// Bounding box .. sort and a force traversal in probe mode on the current
// positions: accelerations go to d.probe, bodies are not advanced, and the
// step counter is put back so the simulation's own stepping is unaffected.
static int probe_pass(nb_sim *s)
{
    nb::Ctl c0;
    int r = read_ctl(s, &c0);
    if (r) return r;
    for (int st = NB_STAGE_BBOX; st <= NB_STAGE_SORT; ++st) {
        r = launch_stage(s, st);
        if (r) return r;
    }
    r = launch_force_method(s, force_method(s), 2);
    if (r) return r;
    HIP_TRY(s, hipMemcpyAsync(&s->d.ctl->step, &c0.step, sizeof(int), hipMemcpyHostToDevice, s->stream));
    return check_device_error(s);
}

static int direct_accel(nb_sim *s, const int32_t *idx, int32_t ns, std::vector<double4> &out)
{
    int r = ensure_check_buffers(s, ns);
    if (r) return r;
    HIP_TRY(s, hipMemcpyAsync(s->chk_idx, idx, (size_t)ns * sizeof(int), hipMemcpyHostToDevice, s->stream));
    hipError_t e = nb::launch_direct64(s->d, s->chk_idx, ns, s->chk_part, s->chk_out, false, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "direct-sum launch: %s", hipGetErrorString(e));
    return fetch(s, out, (const double4 *)s->chk_out, (size_t)ns);
}

int nb_probe_accel(nb_sim *s, float *ax, float *ay, float *az)
{
    if (!s) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    HIP_TRY(s, hipSetDevice(s->device));
    int r = probe_pass(s);
    if (r) return r;
    std::vector<float4> a;
    r = fetch(s, a, (const float4 *)s->d.probe, (size_t)s->d.n);
    if (r) return r;
    for (size_t i = 0; i < a.size(); ++i) {
        if (ax) ax[i] = a[i].x;
        if (ay) ay[i] = a[i].y;
        if (az) az[i] = a[i].z;
    }
    return NB_OK;
}

int nb_direct_accel(nb_sim *s, const int32_t *idx, int32_t ns, double *out3)
{
    if (!s || !idx || !out3 || ns < 0) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    for (int i = 0; i < ns; ++i)
        if (idx[i] < 0 || idx[i] >= s->d.n) return fail(s, NB_E_ARG, "sample index %d out of range", idx[i]);
    if (ns == 0) return NB_OK;
    HIP_TRY(s, hipSetDevice(s->device));
    std::vector<double4> out;
    int r = direct_accel(s, idx, ns, out);
    if (r) return r;
    for (int i = 0; i < ns; ++i) {
        out3[3 * i + 0] = out[i].x;
        out3[3 * i + 1] = out[i].y;
        out3[3 * i + 2] = out[i].z;
    }
    return NB_OK;
}

int nb_force_error(nb_sim *s, int32_t nsamples, uint64_t seed, double *stats4)
{
    if (!s || !stats4 || nsamples < 1) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    HIP_TRY(s, hipSetDevice(s->device));
    const int n = s->d.n;
    const int ns = std::min(nsamples, n);
    // ns distinct bodies: partial Fisher-Yates shuffle driven by splitmix64
    std::vector<int32_t> perm((size_t)n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    uint64_t x = seed;
    for (int i = 0; i < ns; ++i) {
        x += 0x9e3779b97f4a7c15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        const int j = i + (int)(z % (uint64_t)(n - i));
        std::swap(perm[i], perm[j]);
    }
    int r = probe_pass(s);
    if (r) return r;
    std::vector<double4> ref;
    r = direct_accel(s, perm.data(), ns, ref);
    if (r) return r;
    std::vector<float4> a;
    r = fetch(s, a, (const float4 *)s->d.probe, (size_t)n);
    if (r) return r;
    std::vector<double> err((size_t)ns);
    double mean = 0.0;
    for (int i = 0; i < ns; ++i) {
        const float4 g = a[perm[i]];
        const double dx = g.x - ref[i].x, dy = g.y - ref[i].y, dz = g.z - ref[i].z;
        const double den = std::sqrt(ref[i].x * ref[i].x + ref[i].y * ref[i].y + ref[i].z * ref[i].z);
        err[i] = std::sqrt(dx * dx + dy * dy + dz * dz) / std::max(den, 1e-300);
        mean += err[i];
    }
    std::sort(err.begin(), err.end());
    stats4[0] = err[(size_t)(ns - 1) / 2];            // median
    stats4[1] = err[(size_t)((ns - 1) * 0.9)];        // 90th percentile
    stats4[2] = err[(size_t)ns - 1];                  // maximum
    stats4[3] = mean / ns;                            // mean
    s->last_check_median = stats4[0];
    return NB_OK;
}

// ---- learned coarse-theta correction ---------------------------------------

void nb_default_train_params(nb_train_params *tp)
{
    if (!tp) return;
    std::memset(tp, 0, sizeof *tp);
    tp->steps = 50;
    tp->groups_per_step = 512;
    tp->iters = 4000;
    tp->batch = 4096;
    tp->lr = 2e-3f;
    tp->lr_final = 1e-4f;
    tp->beta1 = 0.9f;
    tp->beta2 = 0.999f;
    tp->adam_eps = 1e-8f;
    tp->seed = 1;
}

int nb_set_model(nb_sim *s, const float *model, int32_t count)
{
    if (!s || !model) return NB_E_ARG;
    if (count != NB_MODEL_FLOATS) return fail(s, NB_E_MODEL, "model blob has %d floats, expected %d", count, NB_MODEL_FLOATS);
    HIP_TRY(s, hipSetDevice(s->device));
    return install_model(s, model);
}

int nb_get_model(nb_sim *s, float *model, int32_t count)
{
    if (!s || !model) return NB_E_ARG;
    if (!s->have_model) return fail(s, NB_E_MODEL, "no model installed");
    if (count < NB_MODEL_FLOATS) return fail(s, NB_E_ARG, "model buffer holds %d of %d floats", count, NB_MODEL_FLOATS);
    std::memcpy(model, s->model.data(), sizeof(float) * NB_MODEL_FLOATS);
    return NB_OK;
}

int nb_train(nb_sim *s, const nb_train_params *tp_in, nb_train_report *rep)
{
    if (!s) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    nb_train_params tp;
    nb_default_train_params(&tp);
    if (tp_in) tp = *tp_in;
    if (tp.steps < 1 || tp.groups_per_step < 1 || tp.iters < 0 || tp.batch < 1 || !(tp.lr > 0.0f) ||
        !(tp.lr_final >= 0.0f) || !(tp.beta1 >= 0.0f && tp.beta1 < 1.0f) || !(tp.beta2 >= 0.0f && tp.beta2 < 1.0f) ||
        !(tp.adam_eps > 0.0f))
        return fail(s, NB_E_ARG, "invalid training parameters");
    if (s->p.vote_width != 64) return fail(s, NB_E_ARG, "training needs vote width 64");
    nb::Dev &d = s->d;
    const int ngroups_all = d.n / 64;
    if (ngroups_all < 1) return fail(s, NB_E_ARG, "training needs at least 64 bodies");
    const int G = std::min(tp.groups_per_step, ngroups_all);
    const int rows = G * 64;
    const long long S64 = (long long)tp.steps * rows;
    if (S64 > (1LL << 26)) return fail(s, NB_E_ARG, "%lld training samples exceed 2^26", S64);
    const int S = (int)S64;
    const int B = std::min(tp.batch, S);
    HIP_TRY(s, hipSetDevice(s->device));
    int r = ensure_learn(s);
    if (r) return r;
    const int saved_quad = d.quad;
    d.quad = 1;
    struct QuadRestore {
        nb::Dev &d;
        int v;
        ~QuadRestore() { d.quad = v; }
    } restore{d, saved_quad};

    Scratch m(s);
    nb::Train t{};
    float *X = nullptr, *T = nullptr, *stats = nullptr;
    float4 *frame = nullptr, *afine = nullptr;
    int *gdev = nullptr;
    r = train_buffers(m, t, S, B, tp.iters);
    if (!r) r = m.get(&X, (size_t)S * nb::kFeat);
    if (!r) r = m.get(&T, (size_t)S * nb::kOut);
    if (!r) r = m.get(&stats, (size_t)2 * (nb::kFeat + nb::kOut));
    if (!r) r = m.get(&frame, (size_t)2 * rows);
    if (!r) r = m.get(&afine, (size_t)rows);
    if (!r) r = m.get(&gdev, (size_t)G * tp.steps);
    if (r) return r;
    t.X = X;
    t.T = T;

    // the vote groups of every step: G distinct groups each, drawn up front
    uint64_t rng = tp.seed ^ 0x243f6a8885a308d3ull;
    std::vector<int> perm((size_t)ngroups_all), groups((size_t)G * tp.steps);
    for (int k = 0; k < tp.steps; ++k) {
        for (int i = 0; i < ngroups_all; ++i) perm[i] = i;
        for (int i = 0; i < G; ++i) std::swap(perm[i], perm[i + (int)(splitmix64(rng) % (uint64_t)(ngroups_all - i))]);
        std::sort(perm.begin(), perm.begin() + G);
        std::copy(perm.begin(), perm.begin() + G, groups.begin() + (size_t)k * G);
    }
    r = upload(s, gdev, groups.data(), groups.size() * sizeof(int));
    if (r) return r;

    const auto t0 = std::chrono::steady_clock::now();
    const int fused = (s->p.flags & NB_FUSED_UPDATE) ? 1 : 0;
    for (int k = 0; k < tp.steps; ++k) {
        for (int st = NB_STAGE_BBOX; st <= NB_STAGE_SORT; ++st) {
            r = launch_stage(s, st);
            if (r) return r;
        }
        hipError_t e = nb::launch_coarse(d, s->L, nb::kOutRaw, nullptr, gdev + (size_t)k * G, G,
                                         X + (size_t)k * rows * nb::kFeat, frame, nullptr, 0, s->stream);
        if (e == hipSuccess)
            e = nb::launch_coarse(d, s->L, nb::kOutAccel, dq_plain_ptr(s), gdev + (size_t)k * G, G, nullptr,
                                  nullptr, afine, 0, s->stream);
        if (e == hipSuccess) e = nb::launch_make_target(frame, afine, rows, T + (size_t)k * rows * nb::kOut, s->stream);
        if (e != hipSuccess) return fail(s, NB_E_HIP, "training pair launch: %s", hipGetErrorString(e));
        // the bodies advance with plain theta
        r = launch_force_method(s, NB_MODE_PLAIN, fused);
        if (!r) r = launch_stage(s, NB_STAGE_INTEGRATE);
        if (r) return r;
    }
    r = check_device_error(s);
    if (r) return r;
    const auto t1 = std::chrono::steady_clock::now();

    float *xmu = stats, *xsd = stats + nb::kFeat, *ymu = stats + 2 * nb::kFeat, *ysd = ymu + nb::kOut;
    hipError_t e = nb::launch_standardize(X, S, nb::kFeat, xmu, xsd, s->stream);
    if (e == hipSuccess) e = nb::launch_standardize(T, S, nb::kOut, ymu, ysd, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "standardize launch: %s", hipGetErrorString(e));

    // seeded initialization: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    std::vector<float> P((size_t)nb::kMlpParams);
    uint64_t irng = tp.seed * 0x9e3779b97f4a7c15ull + 7;
    auto uni = [&](float bound) {
        return (float)((double)(splitmix64(irng) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * bound;
    };
    const float bnd1 = 1.0f / std::sqrt((float)nb::kFeat), bnd2 = 1.0f / std::sqrt((float)nb::kHid);
    for (int i = 0; i < nb::kOffB1; ++i) P[i] = uni(bnd1);
    for (int i = nb::kOffB1; i < nb::kOffW2; ++i) P[i] = uni(bnd1);
    for (int i = nb::kOffW2; i < nb::kMlpParams; ++i) P[i] = uni(bnd2);
    HIP_TRY(s, hipMemcpyAsync(t.P, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemsetAsync(t.M, 0, (size_t)nb::kMlpParams * sizeof(float), s->stream));
    HIP_TRY(s, hipMemsetAsync(t.V, 0, (size_t)nb::kMlpParams * sizeof(float), s->stream));
    const uint64_t bseed = tp.seed * 2 + 1; // nonzero: random minibatches
    double p1 = 1.0, p2 = 1.0;
    for (int it = 0; it < tp.iters; ++it) {
        const double frac = tp.iters > 1 ? (double)it / (tp.iters - 1) : 0.0;
        const float lr = (float)(tp.lr_final + 0.5 * (tp.lr - tp.lr_final) * (1.0 + std::cos(M_PI * frac)));
        p1 *= tp.beta1;
        p2 *= tp.beta2;
        e = nb::launch_train_iter(t, B, it, bseed, true, lr, tp.beta1, tp.beta2, tp.adam_eps, (float)(1.0 / (1.0 - p1)),
                                  (float)(1.0 / (1.0 - p2)), s->stream);
        if (e != hipSuccess) return fail(s, NB_E_HIP, "training launch: %s", hipGetErrorString(e));
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    const auto t2 = std::chrono::steady_clock::now();

    std::vector<float> blob((size_t)NB_MODEL_FLOATS), st, lh;
    r = fetch(s, P, (const float *)t.P, (size_t)nb::kMlpParams);
    if (!r) r = fetch(s, st, (const float *)stats, (size_t)2 * (nb::kFeat + nb::kOut));
    if (!r && tp.iters > 0) r = fetch(s, lh, (const float *)t.loss_hist, (size_t)tp.iters);
    if (r) return r;
    std::copy(P.begin(), P.end(), blob.begin());
    std::copy(st.begin(), st.end(), blob.begin() + nb::kMlpParams);
    restore.v = s->p.mode != NB_MODE_PLAIN ? 1 : 0;
    r = install_model(s, blob.data());
    if (r) return r;
    if (rep) {
        std::memset(rep, 0, sizeof *rep);
        rep->samples = S;
        rep->loss_first = lh.empty() ? 0.0f : lh.front();
        rep->loss_last = lh.empty() ? 0.0f : lh.back();
        rep->gen_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        rep->train_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    return NB_OK;
}

int nb_coarse_features(nb_sim *s, float *feat, float *acc3, float *e2)
{
    if (!s) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    if (s->p.vote_width != 64) return fail(s, NB_E_ARG, "the coarse traversal needs vote width 64");
    HIP_TRY(s, hipSetDevice(s->device));
    nb::Dev &d = s->d;
    int r = ensure_quad(s);
    if (r) return r;
    const int saved_quad = d.quad;
    d.quad = 1;
    struct QuadRestore {
        nb::Dev &d;
        int v;
        ~QuadRestore() { d.quad = v; }
    } restore{d, saved_quad};
    const size_t n = (size_t)d.n;
    Scratch m(s);
    float *xr = nullptr;
    float4 *fr = nullptr;
    r = m.get(&xr, n * nb::kFeat);
    if (!r) r = m.get(&fr, 2 * n);
    if (r) return r;
    nb::Ctl c0;
    r = read_ctl(s, &c0);
    if (r) return r;
    for (int st = NB_STAGE_BBOX; st <= NB_STAGE_SORT; ++st) {
        r = launch_stage(s, st);
        if (r) return r;
    }
    hipError_t e = nb::launch_coarse(d, s->L, nb::kOutRaw, nullptr, nullptr, 0, xr, fr, nullptr, 0, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "coarse launch: %s", hipGetErrorString(e));
    HIP_TRY(s, hipMemcpyAsync(&d.ctl->step, &c0.step, sizeof(int), hipMemcpyHostToDevice, s->stream));
    r = check_device_error(s);
    if (r) return r;
    std::vector<float> hx;
    std::vector<float4> hf;
    std::vector<int> srt;
    r = fetch(s, hx, (const float *)xr, n * nb::kFeat);
    if (!r) r = fetch(s, hf, (const float4 *)fr, 2 * n);
    if (!r) r = fetch(s, srt, (const int *)d.sorted, n);
    if (r) return r;
    for (size_t pos = 0; pos < n; ++pos) {
        const size_t b = (size_t)srt[pos];
        if (feat) std::memcpy(feat + b * nb::kFeat, hx.data() + pos * nb::kFeat, nb::kFeat * sizeof(float));
        const float4 a = hf[2 * pos], w = hf[2 * pos + 1];
        if (acc3) {
            acc3[3 * b + 0] = a.x;
            acc3[3 * b + 1] = a.y;
            acc3[3 * b + 2] = a.z;
        }
        if (e2) {
            const float inv = a.w > 0.f ? 1.0f / a.w : 0.f;
            e2[3 * b + 0] = w.x * inv;
            e2[3 * b + 1] = w.y * inv;
            e2[3 * b + 2] = w.z * inv;
        }
    }
    return NB_OK;
}

int nb_mlp_forward(nb_sim *s, const float *x, int32_t n, float *y)
{
    if (!s || n < 0 || (n > 0 && (!x || !y))) return NB_E_ARG;
    if (!s->have_model) return fail(s, NB_E_MODEL, "no model installed");
    if (n == 0) return NB_OK;
    HIP_TRY(s, hipSetDevice(s->device));
    Scratch m(s);
    float *xd = nullptr, *yd = nullptr;
    int r = m.get(&xd, (size_t)n * nb::kFeat);
    if (!r) r = m.get(&yd, (size_t)n * nb::kOut);
    if (r) return r;
    HIP_TRY(s, hipMemcpyAsync(xd, x, (size_t)n * nb::kFeat * sizeof(float), hipMemcpyHostToDevice, s->stream));
    hipError_t e = nb::launch_mlp_rows(s->L, xd, n, yd, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "network launch: %s", hipGetErrorString(e));
    HIP_TRY(s, hipMemcpyAsync(y, yd, (size_t)n * nb::kOut * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}

namespace {
// Full-batch passes of the FP32 training kernels on host data, from the
// installed model's parameters (the test path of the training kernels).
int host_data_passes(nb_sim *s, const float *x, const float *t, int n, int iters, bool adam, float lr, float *loss,
                     float *out)
{
    if (!s || !x || !t || n < 1 || iters < 1) return NB_E_ARG;
    if (!s->have_model) return fail(s, NB_E_MODEL, "no model installed");
    HIP_TRY(s, hipSetDevice(s->device));
    Scratch m(s);
    nb::Train tr{};
    float *xd = nullptr, *td = nullptr;
    int r = train_buffers(m, tr, n, n, iters);
    if (!r) r = m.get(&xd, (size_t)n * nb::kFeat);
    if (!r) r = m.get(&td, (size_t)n * nb::kOut);
    if (r) return r;
    tr.X = xd;
    tr.T = td;
    HIP_TRY(s, hipMemcpyAsync(xd, x, (size_t)n * nb::kFeat * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(td, t, (size_t)n * nb::kOut * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(tr.P, s->model.data(), (size_t)nb::kMlpParams * sizeof(float), hipMemcpyHostToDevice,
                              s->stream));
    HIP_TRY(s, hipMemsetAsync(tr.M, 0, (size_t)nb::kMlpParams * sizeof(float), s->stream));
    HIP_TRY(s, hipMemsetAsync(tr.V, 0, (size_t)nb::kMlpParams * sizeof(float), s->stream));
    double p1 = 1.0, p2 = 1.0;
    for (int it = 0; it < iters; ++it) {
        p1 *= 0.9;
        p2 *= 0.999;
        hipError_t e = nb::launch_train_iter(tr, n, it, 0, adam, lr, 0.9f, 0.999f, 1e-8f, (float)(1.0 / (1.0 - p1)),
                                             (float)(1.0 / (1.0 - p2)), s->stream);
        if (e != hipSuccess) return fail(s, NB_E_HIP, "training launch: %s", hipGetErrorString(e));
    }
    if (loss) HIP_TRY(s, hipMemcpyAsync(loss, tr.loss_hist, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipMemcpyAsync(out, adam ? tr.P : tr.G, (size_t)nb::kMlpParams * sizeof(float), hipMemcpyDeviceToHost,
                              s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    return NB_OK;
}
} // namespace

int nb_mlp_loss_grad(nb_sim *s, const float *x, const float *t, int32_t n, float *loss, float *grad)
{
    if (!grad) return NB_E_ARG;
    return host_data_passes(s, x, t, n, 1, false, 0.0f, loss, grad);
}

int nb_mlp_adam(nb_sim *s, const float *x, const float *t, int32_t n, int32_t iters, float lr, float *params)
{
    if (!params || !(lr > 0.0f)) return NB_E_ARG;
    return host_data_passes(s, x, t, n, iters, true, lr, nullptr, params);
}

// ---- diagnostics and checkpoints --------------------------------------------

int nb_diagnostics(nb_sim *s, int32_t pe_samples, uint64_t seed, double *out)
{
    if (!s || !out) return NB_E_ARG;
    if (!s->have_state) return fail(s, NB_E_STATE, "nb_set_state has not been called");
    HIP_TRY(s, hipSetDevice(s->device));
    const int n = s->d.n;
    Scratch m(s);
    double *mp = nullptr;
    int r = m.get(&mp, (size_t)nb::kMomentBlocks * 11);
    if (r) return r;
    hipError_t e = nb::launch_moments(s->d, mp, s->stream);
    if (e != hipSuccess) return fail(s, NB_E_HIP, "moments launch: %s", hipGetErrorString(e));
    std::vector<double> part;
    r = fetch(s, part, (const double *)mp, (size_t)nb::kMomentBlocks * 11);
    if (r) return r;
    double sum[11] = {};
    for (int b = 0; b < nb::kMomentBlocks; ++b)
        for (int k = 0; k < 11; ++k) sum[k] += part[(size_t)b * 11 + k];
    // potential: targets in batches (the direct sum keeps 64 partial sums per target)
    const bool exact = pe_samples <= 0 || pe_samples >= n;
    const int ns = exact ? n : pe_samples;
    std::vector<int32_t> idx((size_t)ns);
    if (exact) {
        for (int i = 0; i < n; ++i) idx[i] = i;
    } else {
        std::vector<int32_t> perm((size_t)n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        uint64_t x = seed;
        for (int i = 0; i < ns; ++i) std::swap(perm[i], perm[i + (int)(splitmix64(x) % (uint64_t)(n - i))]);
        std::copy(perm.begin(), perm.begin() + ns, idx.begin());
    }
    std::vector<float4> pm;
    r = fetch(s, pm, (const float4 *)s->d.pm4, (size_t)n);
    if (r) return r;
    const int batch = std::min(ns, 65536);
    r = ensure_check_buffers(s, batch);
    if (r) return r;
    double pe = 0.0;
    for (int b0 = 0; b0 < ns; b0 += batch) {
        const int nb_ = std::min(batch, ns - b0);
        std::vector<double4> res;
        HIP_TRY(s, hipMemcpyAsync(s->chk_idx, idx.data() + b0, (size_t)nb_ * sizeof(int), hipMemcpyHostToDevice,
                                  s->stream));
        e = nb::launch_direct64(s->d, s->chk_idx, nb_, s->chk_part, s->chk_out, true, s->stream);
        if (e != hipSuccess) return fail(s, NB_E_HIP, "direct-sum launch: %s", hipGetErrorString(e));
        r = fetch(s, res, (const double4 *)s->chk_out, (size_t)nb_);
        if (r) return r;
        for (int i = 0; i < nb_; ++i) pe += (double)pm[idx[b0 + i]].w * res[i].w;
    }
    pe *= -0.5 * (double)n / (double)ns;
    out[0] = sum[4];
    out[1] = pe;
    out[2] = sum[1];
    out[3] = sum[2];
    out[4] = sum[3];
    out[5] = sum[5];
    out[6] = sum[6];
    out[7] = sum[7];
    const double inv = sum[0] > 0.0 ? 1.0 / sum[0] : 0.0;
    out[8] = sum[8] * inv;
    out[9] = sum[9] * inv;
    out[10] = sum[10] * inv;
    out[11] = sum[0];
    return NB_OK;
}

int nb_get_masses(nb_sim *s, float *mass)
{
    if (!s || !mass) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    std::vector<float4> p;
    int r = fetch(s, p, (const float4 *)s->d.pm4, (size_t)s->d.n);
    if (r) return r;
    for (size_t i = 0; i < p.size(); ++i) mass[i] = p[i].w;
    return NB_OK;
}

int nb_restore_state(nb_sim *s, const double *x, const double *y, const double *z, const float *vx,
                     const float *vy, const float *vz, const float *ax, const float *ay, const float *az,
                     const float *mass, int32_t step, int32_t max_depth)
{
    if (!s) return NB_E_ARG;
    if (!x || !y || !z || !vx || !vy || !vz || !ax || !ay || !az || !mass)
        return fail(s, NB_E_ARG, "null state array");
    if (step < -1 || max_depth < 1 || max_depth > NB_MAX_DEPTH) return fail(s, NB_E_ARG, "invalid step or depth");
    HIP_TRY(s, hipSetDevice(s->device));
    const size_t n = (size_t)s->d.n;
    std::vector<float> fx(n), fy(n), fz(n);
    for (size_t i = 0; i < n; ++i) {
        fx[i] = (float)x[i];
        fy[i] = (float)y[i];
        fz[i] = (float)z[i];
    }
    int r = nb_set_state(s, fx.data(), fy.data(), fz.data(), vx, vy, vz, mass);
    if (r) return r;
    nb::Dev &d = s->d;
    {
        std::vector<float4> acc(n);
        for (size_t i = 0; i < n; ++i) acc[i] = make_float4(ax[i], ay[i], az[i], 0.f);
        r = upload(s, d.acc4, acc.data(), n * sizeof(float4));
        if (r) return r;
    }
    if (d.fp64) {
        std::vector<double4> p64(n);
        for (size_t i = 0; i < n; ++i) p64[i] = make_double4(x[i], y[i], z[i], 0.0);
        r = upload(s, d.pos64, p64.data(), n * sizeof(double4));
        if (r) return r;
    }
    HIP_TRY(s, hipMemcpyAsync(&d.ctl->step, &step, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipMemcpyAsync(&d.ctl->max_depth, &max_depth, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    s->hint_valid = false;
    return NB_OK;
}

int nb_get_counters(nb_sim *s, nb_counters *out)
{
    if (!s || !out) return NB_E_ARG;
    HIP_TRY(s, hipSetDevice(s->device));
    nb::Ctl c;
    int r = read_ctl(s, &c);
    if (r) return r;
    std::memset(out, 0, sizeof *out);
    out->step = c.step;
    out->cells = c.n_cells;
    out->max_depth = c.max_depth;
    out->live_cells = c.live_cells;
    out->large_buckets = c.large_buckets;
    out->interactions = s->last_interactions;
    out->cell_interactions = s->last_cell_interactions;
    out->error = c.error;
    out->fallbacks = s->fallbacks;
    for (int i = 0; i < 8; ++i) out->stage_ms[i] = s->stage_ms[i];
    out->last_check_median = s->last_check_median;
    out->check_ms = s->check_ms;
    out->checks = s->checks;
    out->coarse_alpha = s->alpha_now;
    return NB_OK;
}

int nb_get_control(nb_sim *s, nb_control *out)
{
    if (!s || !out) return NB_E_ARG;
    std::memset(out, 0, sizeof *out);
    out->steps_since_model = s->steps_since_model;
    out->check_seed = s->check_seed;
    out->coarse_alpha = s->alpha_now;
    out->fell_back = s->fell_back ? 1 : 0;
    out->fail_streak = s->fail_streak;
    out->fallbacks = s->fallbacks;
    out->checks = s->checks;
    out->last_check_median = s->last_check_median;
    return NB_OK;
}

int nb_set_control(nb_sim *s, const nb_control *in)
{
    if (!s || !in) return NB_E_ARG;
    if (in->steps_since_model < 0 || !(in->coarse_alpha >= 0.0f) || !std::isfinite(in->coarse_alpha) ||
        (in->fell_back != 0 && in->fell_back != 1) || in->fail_streak < 0 || in->fallbacks < 0 || in->checks < 0)
        return fail(s, NB_E_ARG, "invalid control state");
    s->steps_since_model = in->steps_since_model;
    s->check_seed = in->check_seed;
    s->alpha_now = in->coarse_alpha;
    s->d.alpha2_coarse = in->coarse_alpha > 0.0f ? (float)((double)in->coarse_alpha * (double)in->coarse_alpha) : 0.0f;
    s->fell_back = in->fell_back != 0;
    s->fail_streak = in->fail_streak;
    s->fallbacks = in->fallbacks;
    s->checks = in->checks;
    s->last_check_median = in->last_check_median;
    return NB_OK;
}

int nb_set_timing(nb_sim *s, int32_t enabled)
{
    if (!s) return NB_E_ARG;
    s->timing = enabled != 0;
    return NB_OK;
}

const char *nb_last_error(const nb_sim *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

const char *nb_status_string(int code)
{
    switch (code) {
    case NB_OK: return "ok";
    case NB_E_HIP: return "HIP runtime error";
    case NB_E_POOL: return "octree cell pool exhausted";
    case NB_E_DEPTH: return "octree too deep";
    case NB_E_ARG: return "invalid argument";
    case NB_E_SYNC: return "bounded device wait ran out";
    case NB_E_STATE: return "call out of order";
    case NB_E_STACK: return "traversal stack overflow";
    case NB_E_MODEL: return "learned model missing or malformed";
    default: return "unknown status";
    }
}

} // extern "C"
