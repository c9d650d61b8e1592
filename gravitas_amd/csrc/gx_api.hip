// gx_api.hip -- C ABI and step orchestration (host side of the backend).
//
// One force step enqueues, on one stream and with no host synchronisation in
// between: step init, collect and keys (3 launches), the radix passes (3
// launches each), the sorted copies, the topology (4 launches), the leaf
// moments and one moment launch per internal level, the walk and the hole
// fix-up.  Sizes that only the device knows (stars per instance, nodes,
// walk groups) are read by the kernels from the Meta record and the instance
// arrays; grids are sized from the body count n and the instance count B.  A
// single system is B = 1.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "gravitas_hip.h"
#include "gx_common.h"
#include "gx_kernels.h"

using namespace gx;

namespace {

thread_local char g_err[512];

int set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return set_err(GX_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

std::once_flag g_once;
int g_init_code = GX_ERR_NO_DEVICE;
char g_init_msg[256] = "gx_init not called";
char g_name[256] = "";

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Set while gx_accelerations_noalloc runs: its buffers are slices of a
// static scratch array, and a step that needs more must fail, not allocate.
thread_local bool t_noAlloc = false;

int ensure(DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return GX_OK;
    if (t_noAlloc)
        return set_err(GX_ERR_CAPACITY, "a buffer of %zu bytes exceeds the no-allocation scratch (slice of %zu)", bytes,
                       b.bytes);
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    HIPCHK(hipMalloc(&b.p, want));
    b.bytes = want;
    return GX_OK;
}

constexpr int kPhases = 9;
constexpr int kTreeGrid = 1024;
enum { kNoEvents = 0, kStepEvents = 1, kGraphEvents = 2 };

// The arrays of one step.
struct StepArgs {
    const double *x, *y, *z, *m;
    const uint8_t *kind, *alive;
    long long n;
    int dmOff;
    double *ax, *ay, *az;
    const uint8_t* active;  // walk only these bodies (state order); null: all
};

}  // namespace

struct gx_solver {
    double theta = 0;
    int monopole = 0;
    int walk = GX_WALK_FP32;
    int timing = 0;
    int gsize = 64;           // stars per walk group
    int persistent = 1;       // FP32 walk work queue: 0 never, 1 when groups outnumber resident workgroups, 2 always
    int walkConfig = kWalkDefault;  // stack and staging sizes of the FP32 walk (kWalkConfig)
    int cus = 0;                    // compute units of the device
    int blocksPerCu[kWalkConfigs] = {};  // resident walk workgroups per CU, per configuration
    bool splitWalk = false;   // GX_SPLIT_WALK=1: traversal and evaluation as two kernels
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    // host-API staging
    DevBuf x, y, z, m, kind, alive, ax, ay, az;
    // arena
    DevBuf keys, keys2, order, order2, rank, holes, hist, digitTotal, blockStar, blockHole, blockBox, spos, instOf;
    DevBuf trav, m64, m32, m16, meta, partial;
    DevBuf lfull, lpart, lleaf, lcounts;
    // active sets and block steps
    DevBuf actS, blockAct, tOff, tlist, activeMask, rung, rungHist, diag;
    // split steps near black holes
    DevBuf nearFlag, nearList, tileCnt, splitScalars, farF, fastF;
    // frame dragging: list, chains (increments, two guesses), segment summaries, scalars
    DevBuf dragFlag, dragList, dragD, dragS0, dragS1, dragSig, dragGuess, dragSeg, dragSegStart, dragSegMode,
        dragSpec, dragScal, dragCount;
    int dragIters = 16;  // GX_DRAG_ITERATIONS: fixed-point iterations before the sequential pass (0: always)
    gx_drag_stats dragStats = {};
    unsigned long long* hhist = nullptr;  // pinned mirror of hist
    int cf = 2048, cp = 512, cl = 256;  // FP32-walk list capacities per group (grow on demand)
    long long capNodes = 0;
    // Levels of internal-node moments launched (levels-1 .. 0): the last
    // step's depth + 2 (a deeper tree raises kFlagDepth and the step is
    // repeated with every level); all levels while a graph is captured.
    int levels = kMortonBits;
    bool allLevels = false;
    // instances: device config (offsets | theta | eps2) and per-instance results
    DevBuf instCfg, boxEnc, starCount, starOff, groupOff, root, half;
    char* instHost = nullptr;  // pinned staging of the config
    size_t instHostBytes = 0;
    std::vector<char> instLast;  // last uploaded config
    hipEvent_t instCopied = nullptr;
    int B = 0;
    bool eps0 = false;
    Meta* hmeta = nullptr;  // pinned mirror of the device Meta
    hipEvent_t ev[kPhases + 1] = {};
    hipEvent_t gev[2] = {};  // around a run of graph replays
    int lastEvents = kNoEvents;  // which events time the pending step
    int graphSteps = 0;
    bool capturing = false;  // no event records inside a captured graph
    float phase[kPhases] = {};
    double lastSeconds = 0;
    gx_stats stats = {};
    long long fallbacks = 0;
    bool pending = false;
    int lastWalk = GX_WALK_FP32;
    // captured KDK step
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::vector<void*> graphBufs;  // the arena's buffers when the step was captured
    // autotuner choices per workload (body count, instances, walk, monopole)
    struct Tuned {
        long long n;
        int B, walk, monopole;
        int gsize, persistent, config;
        std::vector<gx_tune_entry> table;
    };
    std::vector<Tuned> tuned;
    bool nodesSet = false;  // gx_solver_reserve set the node capacity: no automatic floor
    // the solver of gx_accelerations_noalloc: list capacities are fixed (an
    // overflow repeats the walk in FP64 instead of growing them), the work
    // queue uses at most maxSlots list regions, no events, no cached uploads
    bool fixedCaps = false;
    int maxSlots = 0;

    // Every device buffer the solver owns.
    std::vector<DevBuf*> buffers() {
        return {&x,        &y,       &z,        &m,          &kind,         &alive,        &ax,        &ay,
                &az,       &keys,    &keys2,    &order,      &order2,       &rank,         &holes,     &hist,
                &digitTotal, &blockStar, &blockHole, &blockBox, &spos,   &instOf,       &trav,         &m64,          &m32,       &m16,
                &meta,     &partial, &lfull,    &lpart,      &lleaf,        &lcounts,      &instCfg,   &boxEnc,
                &starCount, &starOff, &groupOff, &root,      &half,         &actS,         &blockAct,  &tOff,
                &tlist,    &activeMask, &rung,  &rungHist,   &diag,         &nearFlag,     &nearList,  &tileCnt,
                &splitScalars, &farF, &fastF,   &dragFlag,   &dragList,     &dragD,        &dragS0,    &dragS1,
                &dragSig,  &dragGuess, &dragSeg, &dragSegStart, &dragSegMode, &dragSpec,   &dragScal,  &dragCount};
    }
    std::vector<void*> buffer_ptrs() {
        std::vector<void*> p;
        for (DevBuf* b : buffers()) p.push_back(b->p);
        return p;
    }
};

// The holes' histories for retarded gravity (Simulation.histories,
// internal/sim/retarded.go:62-73, 163-209), keyed by body ID: each a ring of
// kHistLen outer steps (x, y, z, vx, vy, vz; 384 KiB) in device memory, its
// header (head, n, t0, appeared, dt) on the host.
struct gx_retard {
    struct Hist {
        uint64_t id;
        double* ring;
        double t0, appeared, dt;
        int n, head;
    };
    std::vector<Hist> hists;
    DevBuf bound;   // HistDev per hole slot of the current step (bindHistories)
    DevBuf pushes;  // HistPush records of one push kernel
    DevBuf ids;     // IDs of the alive holes, slot order
    std::vector<uint64_t> holeIds;
    std::vector<int32_t> holeIdx;
};

extern "C" {

int gx_init(char* err, int errlen) {
    std::call_once(g_once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
            g_init_code = GX_ERR_NO_DEVICE;
            snprintf(g_init_msg, sizeof g_init_msg, "no HIP device");
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            g_init_code = GX_ERR_NO_DEVICE;
            snprintf(g_init_msg, sizeof g_init_msg, "hipGetDeviceProperties failed");
            return;
        }
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            g_init_code = GX_ERR_CODE_OBJECT;
            snprintf(g_init_msg, sizeof g_init_msg, "device %s is not gfx950", prop.gcnArchName);
            return;
        }
        snprintf(g_name, sizeof g_name, "%s (%s)", prop.name, prop.gcnArchName);
        g_init_code = GX_OK;
        g_init_msg[0] = 0;
    });
    if (err && errlen > 0) {
        strncpy(err, g_init_msg, (size_t)errlen);
        err[errlen - 1] = 0;
    }
    return g_init_code;
}

void gx_name(char* out, int len) {
    if (!out || len <= 0) return;
    strncpy(out, g_init_code == GX_OK ? g_name : "", (size_t)len);
    out[len - 1] = 0;
}

const char* gx_error(void) { return g_err; }

}  // extern "C"

// Error reporting and device check for the other translation units (the
// hydro in gx_sph.hip).
int gx::set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int gx::device_ready() {
    if (gx_init(nullptr, 0) != GX_OK) return set_err(g_init_code, "%s", g_init_msg);
    return GX_OK;
}

extern "C" {

gx_solver* gx_solver_new(void) {
    if (gx_init(nullptr, 0) != GX_OK) {
        set_err(g_init_code, "%s", g_init_msg);
        return nullptr;
    }
    gx_solver* s = new gx_solver();
    if (hipStreamCreateWithFlags(&s->own, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&s->hmeta, sizeof(Meta), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->hhist, 64 * 8, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->instCopied, hipEventDisableTiming) != hipSuccess) {
        set_err(GX_ERR_HIP, "stream, event or pinned allocation failed");
        delete s;
        return nullptr;
    }
    memset(s->hmeta, 0, sizeof(Meta));
    s->stream = s->own;
    for (auto& e : s->ev) (void)hipEventCreate(&e);
    for (auto& e : s->gev) (void)hipEventCreate(&e);
    const char* sw = getenv("GX_SPLIT_WALK");
    s->splitWalk = sw && sw[0] == '1';
    const char* di = getenv("GX_DRAG_ITERATIONS");
    if (di && di[0]) s->dragIters = std::max(0, atoi(di));
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        s->cus = 0;
    for (int c = 0; c < kWalkConfigs; c++) s->blocksPerCu[c] = std::max(walk32_blocks_per_cu(c), 1);
    return s;
}

void gx_solver_free(gx_solver* s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->stream);
    if (s->graphExec) (void)hipGraphExecDestroy(s->graphExec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (DevBuf* b : s->buffers())
        if (b->p) (void)hipFree(b->p);
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : s->gev)
        if (e) (void)hipEventDestroy(e);
    if (s->instCopied) (void)hipEventDestroy(s->instCopied);
    if (s->instHost) (void)hipHostFree(s->instHost);
    if (s->hmeta) (void)hipHostFree(s->hmeta);
    if (s->hhist) (void)hipHostFree(s->hhist);
    if (s->own) (void)hipStreamDestroy(s->own);
    delete s;
}

int gx_solver_config(gx_solver* s, double theta, int monopole_only, int walk, int timing) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (!(theta >= 0) || walk < GX_WALK_FP32 || walk > GX_WALK_FP32_Q16)
        return set_err(GX_ERR_ARG, "bad theta %g or walk %d", theta, walk);
    s->theta = theta;
    s->monopole = monopole_only ? 1 : 0;
    s->walk = walk;
    s->timing = timing ? 1 : 0;
    return GX_OK;
}

int gx_solver_tune(gx_solver* s, int group_size, int persistent) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (group_size != 32 && group_size != 64) return set_err(GX_ERR_ARG, "group size %d not in {32, 64}", group_size);
    if (persistent < 0 || persistent > 2) return set_err(GX_ERR_ARG, "persistent %d not in {0, 1, 2}", persistent);
    s->gsize = group_size;
    s->persistent = persistent;
    return GX_OK;
}

int gx_solver_tune_walk(gx_solver* s, int group_size, int persistent, int stack, int parts, int waves) {
    int rc = gx_solver_tune(s, group_size, persistent);
    if (rc) return rc;
    for (int c = 0; c < kWalkConfigs; c++)
        if (kWalkConfig[c].stack == stack && kWalkConfig[c].parts == parts && kWalkConfig[c].waves == waves) {
            s->walkConfig = c;
            return GX_OK;
        }
    return set_err(GX_ERR_ARG,
                   "no walk built with stack %d, staging %d and %d waves (4 waves: stack 384 with staging 128, "
                   "256 or 384, stack 768 with 256 or 384; 5 waves: stack 384 with 128, 256 or 384; 6 waves: "
                   "stack 384 with 256)",
                   stack, parts, waves);
}

int gx_solver_reserve(gx_solver* s, int64_t nodes, int32_t full, int32_t partial, int32_t leaves) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (nodes < 0 || nodes >= (1LL << 28) || full < 0 || partial < 0 || leaves < 0)
        return set_err(GX_ERR_ARG, "bad capacities");
    if (nodes > 0) {
        s->capNodes = nodes;
        s->nodesSet = true;
    }
    if (full > kMaxListCap || partial > kMaxListCap || leaves > kMaxListCap)
        return set_err(GX_ERR_ARG, "list capacities above %d", kMaxListCap);
    if (full > 0) s->cf = full;
    if (partial > 0) s->cp = partial;
    if (leaves > 0) s->cl = leaves;
    return GX_OK;
}

int gx_solver_set_stream(gx_solver* s, void* stream) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    s->stream = stream ? (hipStream_t)stream : s->own;
    return GX_OK;
}

}  // extern "C"

namespace {

double theta_of(double theta) { return theta == 0 ? 0.9 : theta; }  // bh.go:69-74

inline unsigned blocks(long long n, int per) { return (unsigned)std::max(1LL, (n + per - 1) / per); }

// Instance configuration: body offsets (B+1), theta (B), eps^2 (B).  Uploaded
// only when it changes, so repeated steps (and graph replays) reuse it.
int set_instances(gx_solver* s, int B, const int64_t* off, const double* theta, double thetaAll,
                  const double* eps, double epsAll) {
    if (B < 1) return set_err(GX_ERR_ARG, "need at least one instance");
    if (B > 65536) return set_err(GX_ERR_ARG, "at most 65536 instances");
    const size_t bytes = (size_t)(B + 1) * 8 + (size_t)B * 16;
    if (s->instCopied) HIPCHK(hipEventSynchronize(s->instCopied));  // the previous upload has read the staging
    if (s->instHostBytes < bytes) {
        if (s->instHost) HIPCHK(hipHostFree(s->instHost));
        s->instHost = nullptr;
        s->instHostBytes = 0;
        HIPCHK(hipHostMalloc((void**)&s->instHost, bytes, hipHostMallocDefault));
        s->instHostBytes = bytes;
    }
    long long* o = (long long*)s->instHost;
    double* th = (double*)(o + B + 1);
    double* e2 = th + B;
    bool eps0 = false;
    for (int b = 0; b <= B; b++) o[b] = off[b];
    for (int b = 0; b < B; b++) {
        if (o[b + 1] < o[b]) return set_err(GX_ERR_ARG, "instance offsets must ascend");
        th[b] = theta_of(theta ? theta[b] : thetaAll);
        if (!(th[b] >= 0)) return set_err(GX_ERR_ARG, "bad theta of instance %d", b);
        const double e = eps ? eps[b] : epsAll;
        e2[b] = e * e;
        eps0 = eps0 || !((float)e2[b] > 0.0f);
    }
    int rc;
    if ((rc = ensure(s->instCfg, bytes)) || (rc = ensure(s->boxEnc, (size_t)B * 48)) ||
        (rc = ensure(s->starCount, (size_t)B * 8)) || (rc = ensure(s->starOff, (size_t)(B + 1) * 8)) ||
        (rc = ensure(s->groupOff, (size_t)(B + 1) * 8)) || (rc = ensure(s->root, (size_t)B * 4)) ||
        (rc = ensure(s->half, (size_t)B * 8)))
        return rc;
    if (!s->instCopied) {  // no upload cache (gx_accelerations_noalloc): copy and wait
        HIPCHK(hipMemcpyAsync(s->instCfg.p, s->instHost, bytes, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    } else if (s->B != B || s->instLast.size() != bytes || memcmp(s->instLast.data(), s->instHost, bytes) != 0) {
        HIPCHK(hipMemcpyAsync(s->instCfg.p, s->instHost, bytes, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipEventRecord(s->instCopied, s->stream));
        s->instLast.assign(s->instHost, s->instHost + bytes);
    }
    s->B = B;
    s->eps0 = eps0;
    return GX_OK;
}

Batch make_batch(gx_solver* s) {
    Batch b;
    b.B = s->B;
    b.gsize = s->gsize;
    b.bodyOff = (const long long*)s->instCfg.p;
    b.theta = (const double*)((const long long*)s->instCfg.p + s->B + 1);
    b.eps2 = b.theta + s->B;
    b.boxEnc = (unsigned long long*)s->boxEnc.p;
    b.starCount = (unsigned long long*)s->starCount.p;
    b.starOff = (long long*)s->starOff.p;
    b.groupOff = (long long*)s->groupOff.p;
    b.root = (int*)s->root.p;
    b.half = (double*)s->half.p;
    b.tOff = b.starOff;
    b.tlist = nullptr;
    b.instOf = s->B > 1 ? (const uint16_t*)s->instOf.p : nullptr;
    return b;
}

int reserve(gx_solver* s, long long n) {
    const long long nb = (n + kCollectTile - 1) / kCollectTile;
    const long long nt = (n + kSortTile - 1) / kSortTile;
    int rc;
    if ((rc = ensure(s->keys, n * 8)) || (rc = ensure(s->keys2, n * 8)) || (rc = ensure(s->order, n * 4)) ||
        (rc = ensure(s->order2, n * 4)) || (rc = ensure(s->rank, std::max(n, 1LL) * 4)) ||
        (rc = ensure(s->holes, n * 4)) ||
        (rc = ensure(s->hist, (size_t)256 * std::max(nt, 1LL) * 4)) || (rc = ensure(s->digitTotal, 256 * 4)) ||
        (rc = ensure(s->blockStar, std::max(nb, 1LL) * 4)) || (rc = ensure(s->blockHole, std::max(nb, 1LL) * 4)) ||
        (rc = ensure(s->blockBox, std::max(nb, 1LL) * 48)) ||
        (rc = ensure(s->spos, std::max(n, 1LL) * 32)) || (s->B > 1 && (rc = ensure(s->instOf, n * 2))))
        return rc;
    if (!s->meta.p) {  // allocated once; every field starts at zero (the sticky word included)
        if ((rc = ensure(s->meta, sizeof(Meta)))) return rc;
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, (unsigned long long*)s->meta.p,
                           (int)(sizeof(Meta) / 8), 0ull);
        HIPCHK(hipGetLastError());
    }
    const long long wantNodes = s->nodesSet ? 1 : std::max(4096LL, n);
    if (s->capNodes < wantNodes) s->capNodes = wantNodes;
    if (s->capNodes >= (1LL << 28)) return set_err(GX_ERR_CAPACITY, "more than 2^28 nodes");
    if ((rc = ensure(s->trav, (size_t)s->capNodes * sizeof(TravNode))) ||
        (rc = ensure(s->m64, (size_t)s->capNodes * sizeof(Mult64))) ||
        (rc = ensure(s->m32, (size_t)s->capNodes * sizeof(Mult32))) ||
        (rc = ensure(s->m16, (size_t)s->capNodes * sizeof(Mult16))))
        return rc;
    // ensure() may have over-allocated: use all of it (node indices stay below
    // 2^28), unless the caller set the capacity (gx_solver_reserve)
    if (!s->nodesSet)
        s->capNodes = (long long)std::min({s->trav.bytes / sizeof(TravNode), s->m64.bytes / sizeof(Mult64),
                                           s->m32.bytes / sizeof(Mult32), s->m16.bytes / sizeof(Mult16)});
    s->capNodes = std::min(s->capNodes, (1LL << 28) - 1);
    return GX_OK;
}

// Phase events: with timing on, between the phases.  The first and the last
// (the whole step, for gx_last_seconds) are recorded on every eager step.
void mark(gx_solver* s, int i) {
    if ((s->timing || i == 0 || i == kPhases) && !s->capturing) (void)hipEventRecord(s->ev[i], s->stream);
}

// collect (build.go:62-113): stars in state order into order[] (order2[]
// with spare, for an odd number of sort passes), alive holes into holes[],
// counts and boxes into Meta and the instance arrays.  Needs reserve() first.
void enqueue_collect(gx_solver* s, const StepArgs& a, bool spare) {
    const long long n = a.n;
    const long long nb = (n + kCollectTile - 1) / kCollectTile;
    hipStream_t st = s->stream;
    Meta* meta = (Meta*)s->meta.p;
    const Batch batch = make_batch(s);
    hipLaunchKernelGGL(k_step_init, dim3(blocks(s->B, 256)), dim3(256), 0, st, meta, batch);
    hipLaunchKernelGGL(k_collect_count, dim3(blocks(n, kCollectTile)), dim3(256), 0, st, a.x, a.y, a.z, a.kind,
                       a.alive, n, (int*)s->blockStar.p, (int*)s->blockHole.p, (double*)s->blockBox.p, batch);
    hipLaunchKernelGGL(k_collect_scan, dim3(1), dim3(1024), 0, st, (int)std::max(nb, 1LL), (int*)s->blockStar.p,
                       (int*)s->blockHole.p, (const double*)s->blockBox.p, meta, batch);
    hipLaunchKernelGGL(k_collect_scatter, dim3(blocks(n, kCollectTile)), dim3(256), 0, st, a.x, a.y, a.z, a.kind,
                       a.alive, n, (const int*)s->blockStar.p, (const int*)s->blockHole.p, batch,
                       (int32_t*)(spare ? s->order2.p : s->order.p), (int32_t*)s->holes.p,
                       (uint64_t*)(spare ? s->keys2.p : s->keys.p), (uint16_t*)s->instOf.p);
}

// Build the trees of all instances (collect .. moments).  theta is baked into
// crit2 (build.go:209-211).
int enqueue_build(gx_solver* s, const StepArgs& a) {
    const long long n = a.n;
    int rc = reserve(s, n);
    if (rc) return rc;
    hipStream_t st = s->stream;
    Meta* meta = (Meta*)s->meta.p;
    const Batch batch = make_batch(s);
    const int ntiles = std::max(1, (int)((n + kSortTile - 1) / kSortTile));
    // 8 passes over the 63-bit keys, then (several instances) 1 or 2 passes
    // over the instance id (up to 256 instances, up to 65,536), most
    // significant last; the collect writes into the spare arrays when the
    // count is odd, so that the sorted pairs end in keys/order
    const int passes = 8 + (s->B > 256 ? 2 : s->B > 1 ? 1 : 0);
    const bool odd = passes & 1;
    mark(s, 0);
    enqueue_collect(s, a, odd);
    mark(s, 1);  // the keys are written by the collect scatter (phase "keys" stays empty)
    mark(s, 2);
    uint64_t* kin = (uint64_t*)(odd ? s->keys2.p : s->keys.p);
    uint64_t* kout = (uint64_t*)(odd ? s->keys.p : s->keys2.p);
    int32_t* vin = (int32_t*)(odd ? s->order2.p : s->order.p);
    int32_t* vout = (int32_t*)(odd ? s->order.p : s->order2.p);
    for (int pass = 0; pass < passes; pass++) {
        const int inst = pass >= 8;
        const int shift = inst ? 8 * (pass - 8) : 8 * pass;
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(256), 0, st, kin, vin, meta, shift, inst, batch,
                           (uint32_t*)s->hist.p, ntiles);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kRowScanThreads), 0, st, (uint32_t*)s->hist.p, ntiles,
                           (uint32_t*)s->digitTotal.p);
        hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(256), 0, st, kin, vin, kout, vout, meta, shift, inst,
                           batch, (const uint32_t*)s->hist.p, (const uint32_t*)s->digitTotal.p, ntiles,
                           pass == passes - 1 ? (int32_t*)s->rank.p : (int32_t*)nullptr);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    mark(s, 3);
    hipLaunchKernelGGL(k_scatter_spos, dim3((unsigned)std::min<long long>(blocks(n, 256), 8192)), dim3(256), 0, st,
                       a.x, a.y, a.z, a.m, a.kind, a.alive, n, (const int32_t*)s->rank.p, a.dmOff,
                       (double4*)s->spos.p);
    mark(s, 4);
    // the whole topology at any depth: chains and per-level counts per tile,
    // their scan per level, node records, then ends and links (the sort's
    // spare key and value arrays hold the chains and each chain's first node)
    static_assert(kMaxLevels * ((kSortTile + kTopoTile - 1) / kTopoTile) <= 256,
                  "the per-level counts of the topology tiles fit in the sort's histogram");
    const int ttiles = std::max(1, (int)((n + kTopoTile - 1) / kTopoTile));
    hipLaunchKernelGGL(k_topo_count, dim3(ttiles), dim3(256), 0, st, (const uint64_t*)s->keys.p, meta, batch,
                       (uint16_t*)s->keys2.p, (uint32_t*)s->hist.p, ttiles);
    hipLaunchKernelGGL(k_sort_scan, dim3(kMaxLevels), dim3(kRowScanThreads), 0, st, (uint32_t*)s->hist.p, ttiles,
                       (uint32_t*)s->digitTotal.p);
    hipLaunchKernelGGL(k_topo_nodes, dim3(ttiles), dim3(256), 0, st, (const uint64_t*)s->keys.p,
                       (const uint16_t*)s->keys2.p, (const uint32_t*)s->hist.p, (const uint32_t*)s->digitTotal.p, meta,
                       batch, (TravNode*)s->trav.p, (int32_t*)s->order2.p, s->capNodes, ttiles);
    hipLaunchKernelGGL(k_topo_links, dim3(kTreeGrid), dim3(256), 0, st, (const uint64_t*)s->keys.p,
                       (const uint16_t*)s->keys2.p, (TravNode*)s->trav.p, (const int32_t*)s->order2.p, meta, batch);
    mark(s, 5);
    // all leaves first (and the node count), then the internal nodes level by
    // level, bottom up, from level Lmax-1 (the last step's depth + 1)
    const int Lmax = s->allLevels ? kMortonBits : s->levels;
    hipLaunchKernelGGL(k_moments_leaves, dim3(kTreeGrid), dim3(256), 0, st, (const double4*)s->spos.p,
                       (TravNode*)s->trav.p, (Mult64*)s->m64.p, (Mult32*)s->m32.p, (Mult16*)s->m16.p, meta, batch,
                       Lmax);
    for (int L = Lmax - 1; L >= 0; L--)
        hipLaunchKernelGGL(k_moments_level, dim3(kTreeGrid), dim3(256), 0, st, L, (const double4*)s->spos.p,
                           (TravNode*)s->trav.p, (Mult64*)s->m64.p, (Mult32*)s->m32.p, (Mult16*)s->m16.p, meta,
                           batch);
    mark(s, 6);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

// Target list of an active set (after the build): active sorted stars,
// per-instance target offsets and walk groups.
int enqueue_active(gx_solver* s, const StepArgs& a) {
    const long long n = a.n;
    const long long nb = std::max(1LL, (n + kCollectTile - 1) / kCollectTile);
    int rc;
    if ((rc = ensure(s->actS, std::max(n, 1LL))) || (rc = ensure(s->blockAct, nb * 4)) ||
        (rc = ensure(s->tOff, (size_t)(s->B + 1) * 8)) || (rc = ensure(s->tlist, std::max(n, 1LL) * 4)))
        return rc;
    hipStream_t st = s->stream;
    Meta* meta = (Meta*)s->meta.p;
    const Batch batch = make_batch(s);
    hipLaunchKernelGGL(k_active_count, dim3((unsigned)nb), dim3(256), 0, st, a.active, (const int32_t*)s->order.p,
                       meta, (uint8_t*)s->actS.p, (int*)s->blockAct.p);
    hipLaunchKernelGGL(k_active_scan, dim3(1), dim3(1024), 0, st, (int)nb, (int*)s->blockAct.p, meta);
    hipLaunchKernelGGL(k_active_inst, dim3(blocks((long long)(s->B + 1) * kWave, 256)), dim3(256), 0, st,
                       (const uint8_t*)s->actS.p, (const int*)s->blockAct.p, meta, batch, (long long*)s->tOff.p);
    hipLaunchKernelGGL(k_active_groups, dim3(1), dim3(1024), 0, st, batch, (const long long*)s->tOff.p);
    hipLaunchKernelGGL(k_active_scatter, dim3((unsigned)nb), dim3(256), 0, st, (const uint8_t*)s->actS.p,
                       (const int*)s->blockAct.p, meta, (int32_t*)s->tlist.p);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

// accumulate: the walk ends a step, so its flags go into Meta::sticky (not
// for the FP64 walk that repeats an overflowed FP32 walk's work).
int enqueue_walk(gx_solver* s, int walk, const StepArgs& a, bool accumulate = true) {
    hipStream_t st = s->stream;
    Meta* meta = (Meta*)s->meta.p;
    Batch batch = make_batch(s);
    if (a.active) {
        batch.tOff = (const long long*)s->tOff.p;
        batch.tlist = (const int32_t*)s->tlist.p;
    }
    const long long n = a.n;
    if (walk == GX_WALK_FP32 || walk == GX_WALK_FP32_Q16) {
        const long long ng = (n + s->gsize - 1) / s->gsize + s->B;
        const int cfg = s->walkConfig;
        long long resident = (long long)std::max(s->cus, 1) * s->blocksPerCu[cfg];
        if (s->maxSlots > 0) resident = std::min<long long>(resident, s->maxSlots);
        // the work queue pays once there are more groups than resident
        // workgroups; below that, one workgroup per group (persistent 1).
        // persistent 2 always takes the queue.
        const bool queue = !s->splitWalk && (s->persistent == 2 || (s->persistent == 1 && ng > resident));
        // one list region per resident workgroup with the queue, else per group
        const long long regions = queue ? resident : ng;
        int rc;
        if ((rc = ensure(s->lfull, (size_t)regions * s->cf * 4)) ||
            (rc = ensure(s->lpart, (size_t)regions * s->cp * 12)) ||
            (rc = ensure(s->lleaf, (size_t)regions * s->cl * 12)) ||
            (rc = ensure(s->lcounts, (size_t)regions * sizeof(GroupCounts))))
            return rc;
        Lists lists;
        lists.full = (uint32_t*)s->lfull.p;
        lists.part = (uint3*)s->lpart.p;
        lists.leaf = (uint3*)s->lleaf.p;
        lists.counts = (GroupCounts*)s->lcounts.p;
        lists.cf = s->cf;
        lists.cp = s->cp;
        lists.cl = s->cl;
        lists.perSlot = queue ? 1 : 0;
        launch_walk32((const double4*)s->spos.p, (const TravNode*)s->trav.p, (const Mult32*)s->m32.p,
                      walk == GX_WALK_FP32_Q16 ? (const Mult16*)s->m16.p : nullptr, (const int32_t*)s->order.p, meta,
                      batch, ng, s->eps0, !s->monopole, lists, a.ax, a.ay, a.az, st,
                      s->timing ? s->ev[7] : nullptr, s->splitWalk, queue ? (int)resident : 0, cfg);
    } else {
        hipLaunchKernelGGL(k_walk64, dim3(blocks(n, 256)), dim3(256), 0, st, (const double4*)s->spos.p,
                           (const TravNode*)s->trav.p, (const Mult64*)s->m64.p, (const int32_t*)s->order.p, meta,
                           batch, !s->monopole, a.ax, a.ay, a.az);
    }
    hipLaunchKernelGGL(k_zero_holes, dim3(64), dim3(256), 0, st, (const int32_t*)s->holes.p, meta, a.ax, a.ay, a.az,
                       accumulate ? 1 : 0);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

int enqueue_step(gx_solver* s, const StepArgs& a) {
    if (a.n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (a.n >= (1LL << 31) - 1) return set_err(GX_ERR_ARG, "n = %lld exceeds the int32 index range", a.n);
    int rc = enqueue_build(s, a);
    if (rc) return rc;
    if (a.active && (rc = enqueue_active(s, a))) return rc;
    if ((rc = enqueue_walk(s, s->walk, a))) return rc;
    mark(s, kPhases);
    HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
    s->lastWalk = s->walk;
    s->pending = true;
    s->lastEvents = s->capturing ? kNoEvents : kStepEvents;
    return GX_OK;
}

// Waits for the pending step; fills stats and phase times.
int finish_step(gx_solver* s) {
    if (!s->pending) return GX_OK;
    s->pending = false;
    HIPCHK(hipStreamSynchronize(s->stream));
    const Meta& h = *s->hmeta;
    // GPU time of the step (LastGPUSeconds, gpu_darwin.go:61,155): the first
    // and last events of an eager step, or those around a run of graph
    // replays (the mean per replay)
    {
        float total = 0;
        const hipEvent_t a = s->lastEvents == kGraphEvents ? s->gev[0] : s->ev[0];
        const hipEvent_t b = s->lastEvents == kGraphEvents ? s->gev[1] : s->ev[kPhases];
        if (s->lastEvents != kNoEvents && hipEventElapsedTime(&total, a, b) == hipSuccess)
            s->lastSeconds = total * 1e-3 / (s->lastEvents == kGraphEvents ? std::max(1, s->graphSteps) : 1);
        s->lastEvents = kNoEvents;
    }
    if (s->timing) {
        float total = 0;
        (void)hipEventElapsedTime(&total, s->ev[0], s->ev[kPhases]);
        for (int i = 0; i < 6; i++) (void)hipEventElapsedTime(&s->phase[i], s->ev[i], s->ev[i + 1]);
        (void)hipEventElapsedTime(&s->phase[6], s->ev[6], s->ev[kPhases]);
        s->phase[7] = total;
        s->phase[8] = 0;
        if (s->lastWalk != GX_WALK_FP64 && s->splitWalk) {  // traversal | evaluation split
            float trav = 0;
            (void)hipEventElapsedTime(&trav, s->ev[6], s->ev[7]);
            s->phase[8] = s->phase[6] - trav;
            s->phase[6] = trav;
        }
    }
    int depth = 0;
    for (int l = 0; l < kMaxLevels; l++)
        if (h.levelCount[l] > 0) depth = l;
    s->stats.n_stars = h.nstars;
    s->stats.n_holes = h.nholes;
    s->stats.n_nodes = h.nnodes;
    s->stats.pc_interactions = (int64_t)h.pc;
    // self pairs are counted by the kernels; the reference skips them
    s->stats.pp_interactions = (int64_t)h.pp - (int64_t)h.ntargets;
    s->stats.n_targets = h.ntargets;
    s->stats.max_stack = h.maxStack;
    // this step's flags and those of earlier steps (graph replays) not yet reported
    const int flags = h.flags | (int)h.sticky;
    s->stats.flags = flags;
    if (h.sticky) {  // reported now: clear it on the device (a kernel, never a captured memset)
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, &((Meta*)s->meta.p)->sticky, 1, 0ull);
        HIPCHK(hipGetLastError());
    }
    s->stats.depth = depth;
    s->stats.fallbacks = s->fallbacks;
    s->stats.list_entries = (int64_t)h.listEntries;
    s->stats.rounds = (int64_t)h.rounds;
    s->stats.mixed_nodes = (int64_t)h.mixed;
    for (int i = 0; i < 3; i++) s->stats.max_list[i] = h.maxList[i];
    s->stats.max_list[3] = 0;
    if (flags & kFlagNodeOverflow)
        return set_err(GX_ERR_CAPACITY, "node arena of %lld nodes too small", s->capNodes);
    if (flags & kFlagDepth) return set_err(GX_ERR_CAPACITY, "tree deeper than the levels built");
    if (flags & kFlagTopology)
        return set_err(GX_ERR_PIPELINE, "tree links failed their check (0x%x at node %lld of %lld; flags 0x%x sticky 0x%llx)",
                       h.topoWhy, h.topoNode, h.nnodes, h.flags, h.sticky);
    s->levels = std::min(kMortonBits, depth + 2);
    return GX_OK;
}

// Grows the capacity that overflowed; true if the step must be repeated.
bool regrow(gx_solver* s) {
    const Meta& h = *s->hmeta;
    if (h.flags & kFlagNodeOverflow) {
        s->capNodes *= 2;
        return true;
    }
    if (h.flags & kFlagDepth) {
        s->levels = kMortonBits;
        return true;
    }
    if ((h.flags & kFlagListOverflow) && !s->fixedCaps) {
        // list entries are addressed with 24-bit products in the walk
        auto grown = [](int cap, int seen) {
            return (int)std::min<long long>(kMaxListCap, std::max<long long>(cap, (long long)(seen * 1.25) + 64));
        };
        const int cf = grown(s->cf, h.maxList[0]), cp = grown(s->cp, h.maxList[1]), cl = grown(s->cl, h.maxList[2]);
        if (cf == s->cf && cp == s->cp && cl == s->cl) return false;  // at the limit: the FP64 walk takes over
        s->cf = cf;
        s->cp = cp;
        s->cl = cl;
        return true;
    }
    return false;
}

int run_dev(gx_solver* s, const StepArgs& a) {
    if (s->pending) (void)finish_step(s);
    for (int attempt = 0; attempt < 8; attempt++) {
        int rc = enqueue_step(s, a);
        if (rc) return rc;
        rc = finish_step(s);
        if (rc && rc != GX_ERR_CAPACITY) return rc;
        if (regrow(s)) continue;
        if (s->hmeta->flags & (kFlagStackOverflow | kFlagListOverflow)) {
            // the FP32 walk's LDS stack (or a fixed list) was too small for this
            // tree: redo the walk with the FP64 per-particle walk (same
            // interaction set)
            s->fallbacks++;
            if ((rc = enqueue_walk(s, GX_WALK_FP64, a, false))) return rc;
            HIPCHK(hipStreamSynchronize(s->stream));
            s->stats.fallbacks = s->fallbacks;
        }
        return GX_OK;
    }
    return set_err(GX_ERR_CAPACITY, "capacities did not fit after regrowing");
}

StepArgs step_args(const double* x, const double* y, const double* z, const double* m, const uint8_t* kind,
                   const uint8_t* alive, long long n, int dmOff, double* ax, double* ay, double* az) {
    StepArgs a;
    a.x = x; a.y = y; a.z = z; a.m = m;
    a.kind = kind; a.alive = alive;
    a.n = n;
    a.dmOff = dmOff;
    a.ax = ax; a.ay = ay; a.az = az;
    a.active = nullptr;
    return a;
}

int check_arrays(const StepArgs& a) {
    if (a.n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (a.n > 0 && (!a.x || !a.y || !a.z || !a.m || !a.kind || !a.alive || !a.ax || !a.ay || !a.az))
        return set_err(GX_ERR_ARG, "null array");
    return GX_OK;
}

int set_single(gx_solver* s, long long n, double theta, double eps) {
    const int64_t off[2] = {0, n};
    return set_instances(s, 1, off, nullptr, theta, nullptr, eps);
}

int check_batch(const int64_t* offsets, int32_t count, const double* theta, const double* eps) {
    if (!offsets || !theta || !eps) return set_err(GX_ERR_ARG, "null argument");
    if (count < 1 || offsets[0] != 0) return set_err(GX_ERR_ARG, "bad instance offsets");
    return GX_OK;
}

int kick(gx_solver* s, double* vx, double* vy, double* vz, const double* ax, const double* ay, const double* az,
         long long n, double h) {
    if (n <= 0) return GX_OK;
    hipLaunchKernelGGL(k_kick, dim3((unsigned)std::min<long long>(blocks(n, 256), 4096)), dim3(256), 0, s->stream, n,
                       vx, vy, vz, ax, ay, az, h);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

int drift(gx_solver* s, double* x, double* y, double* z, const double* vx, const double* vy, const double* vz,
          long long n, double dt) {
    if (n <= 0) return GX_OK;
    hipLaunchKernelGGL(k_drift, dim3((unsigned)std::min<long long>(blocks(n, 256), 4096)), dim3(256), 0, s->stream,
                       n, x, y, z, vx, vy, vz, dt);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

// KDK step (sim.go:262-281): forces first if they are not valid, then
// kick, drift, new forces, kick, over every body.
int kdk(gx_solver* s, const StepArgs& a, double* vx, double* vy, double* vz, double dt, int accValid) {
    int rc;
    double* x = (double*)a.x;
    double* y = (double*)a.y;
    double* z = (double*)a.z;
    if (!accValid && (rc = run_dev(s, a))) return rc;
    if ((rc = kick(s, vx, vy, vz, a.ax, a.ay, a.az, a.n, 0.5 * dt)) || (rc = drift(s, x, y, z, vx, vy, vz, a.n, dt)) ||
        (rc = run_dev(s, a)) || (rc = kick(s, vx, vy, vz, a.ax, a.ay, a.az, a.n, 0.5 * dt)))
        return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

// Block step of length T (blockstep.go's event loop, :130-198, for gravity):
// rungs and opening kicks at tick 0; then, from event to event, every body
// drifts, the bodies whose step ends are evaluated (tree of all stars, walk
// of the active ones), kicked, and given a new step.  At the last tick every
// body is active, so the step ends synchronised with valid forces.
int block_step(gx_solver* s, StepArgs a, double* vx, double* vy, double* vz, double T, int levels, double eta,
               double eps, int accValid, gx_block_stats* out) {
    if (levels < 1 || levels > GX_BLOCK_MAX_LEVELS) return set_err(GX_ERR_ARG, "levels %d not in [1, %d]", levels, GX_BLOCK_MAX_LEVELS);
    if (!(T > 0)) return set_err(GX_ERR_ARG, "bad step %g", T);
    const long long n = a.n;
    const int maxRung = levels - 1, full = 1 << maxRung;
    const double unit = T / full;
    const double crit = 2 * eta * eps;
    double* x = (double*)a.x;
    double* y = (double*)a.y;
    double* z = (double*)a.z;
    gx_block_stats st;
    memset(&st, 0, sizeof st);
    st.levels = levels;
    st.bodies = n;
    int rc;
    if ((rc = ensure(s->rung, std::max(n, 1LL))) || (rc = ensure(s->activeMask, std::max(n, 1LL))) ||
        (rc = ensure(s->rungHist, 64 * 8)))
        return rc;
    if (!accValid && (rc = run_dev(s, a))) return rc;
    hipStream_t strm = s->stream;
    unsigned long long* hist = (unsigned long long*)s->rungHist.p;
    const unsigned grid = (unsigned)std::min<long long>(blocks(n, 256), 4096);
    auto kick_at = [&](int now) -> int {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, strm, hist, 64, 0ull);
        if (n > 0)
            hipLaunchKernelGGL(k_block_kick, dim3(grid), dim3(256), 0, strm, n, vx, vy, vz, (const double*)a.ax,
                               (const double*)a.ay, (const double*)a.az, a.kind, a.alive, (int8_t*)s->rung.p, now,
                               full, maxRung, T, crit, hist);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(s->hhist, hist, 64 * 8, hipMemcpyDeviceToHost, strm));
        HIPCHK(hipStreamSynchronize(strm));
        return GX_OK;
    };
    if ((rc = kick_at(0))) return rc;
    for (int r = 0; r <= maxRung; r++) st.rung_count[r] = (int64_t)s->hhist[r];
    StepArgs act = a;
    act.active = (const uint8_t*)s->activeMask.p;
    for (int now = 0; now < full;) {
        int next = full;
        for (int r = 0; r <= maxRung; r++)
            if (s->hhist[r]) next = std::min(next, (now / (full >> r) + 1) * (full >> r));
        if ((rc = drift(s, x, y, z, vx, vy, vz, n, (double)(next - now) * unit))) return rc;
        now = next;
        if (n > 0)
            hipLaunchKernelGGL(k_block_active, dim3(grid), dim3(256), 0, strm, n, (const int8_t*)s->rung.p, now, full,
                               (uint8_t*)s->activeMask.p);
        HIPCHK(hipGetLastError());
        if ((rc = run_dev(s, act))) return rc;
        if (st.events < GX_BLOCK_MAX_TICKS) {
            st.tick[st.events] = now;
            st.active[st.events] = s->hmeta->ntargets;
        }
        st.events++;
        st.targets += s->hmeta->ntargets;
        if ((rc = kick_at(now))) return rc;
    }
    if (out) *out = st;
    return GX_OK;
}

// Frame dragging (internal/sim/framedrag.go).  Scalars in dragScal: the
// frame-dragging flag (int in 8 bytes), the changed flag and the
// element-by-element segment counter (8 bytes each, cleared together), one
// spare slot, then the chain starts and ends startJ | endJ | startV | endV
// (C = 3 nh doubles each).
struct DragBufs {
    double *startJ, *endJ, *startV, *endV;
    int* on;
    unsigned long long *changed, *slow;
};

int drag_bufs(gx_solver* s, long long nh, DragBufs& b) {
    const long long C = 3 * std::max(nh, 1LL);
    int rc = ensure(s->dragScal, (size_t)C * 32 + 32);
    if (rc) return rc;
    if (!s->dragCount.p) {  // iterations and sequential passes of k_drag_small
        if ((rc = ensure(s->dragCount, 16))) return rc;
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, (unsigned long long*)s->dragCount.p, 2, 0ULL);
    }
    unsigned long long* u = (unsigned long long*)s->dragScal.p;
    b.on = (int*)u;
    b.changed = u + 1;
    b.slow = u + 2;
    double* d = (double*)(u + 4);
    b.startJ = d;
    b.endJ = d + C;
    b.startV = d + 2 * C;
    b.endV = d + 3 * C;
    return GX_OK;
}

// Ordered scan of the C chains of D (chain-major, ne per chain) from start
// (gx_drag.hip).  Exact: end receives the chain ends; with S given also the
// value every position sees (segments speculated from Sold's values), and
// *changed is set if any differs from Sold (from the start values if Sold is
// null).  Not exact (S given): S receives an approximation (segments added
// from approximate starts), the guess of the first fixed-point iterations.
int drag_scan(gx_solver* s, const double* D, long long ne, int C, const double* start, const double* Sold, double* S,
              double* end, const DragBufs& b, bool exact) {
    const int nseg = (int)((ne + kDragSeg - 1) / kDragSeg);
    hipStream_t st = s->stream;
    const unsigned wgrid = blocks((long long)C * nseg * 64, 256);
    const unsigned rgrid = blocks((long long)C * nseg * kDragRows, 256);
    double* specStart = (double*)s->dragSpec.p;
    double* specEnd = specStart + (long long)C * nseg;
    int* specOk = (int*)(specEnd + (long long)C * nseg);
    double* rowSig = (double*)s->dragSig.p + (long long)C * nseg;
    hipLaunchKernelGGL(k_scan_segsum, dim3(wgrid), dim3(256), 0, st, D, ne, C, nseg, (double*)s->dragSig.p, rowSig);
    hipLaunchKernelGGL(k_scan_guess, dim3((unsigned)C), dim3(1024), 0, st, (const double*)s->dragSig.p, nseg, start,
                       (double*)s->dragGuess.p);
    if (!exact) {
        hipLaunchKernelGGL(k_scan_spec, dim3(rgrid), dim3(256), 0, st, D, ne, C, nseg, start,
                           (const double*)s->dragGuess.p, (const double*)rowSig, (const double*)nullptr, S,
                           (double*)nullptr, (double*)nullptr, (int*)nullptr);
        HIPCHK(hipGetLastError());
        return GX_OK;
    }
    hipLaunchKernelGGL(k_scan_summary, dim3(wgrid), dim3(256), 0, st, D, ne, C, nseg, (const double*)s->dragGuess.p,
                       (DragSeg*)s->dragSeg.p);
    const bool spec = S && Sold;  // speculate from the previous guess
    if (spec)
        hipLaunchKernelGGL(k_scan_spec, dim3(rgrid), dim3(256), 0, st, D, ne, C, nseg, start,
                           (const double*)s->dragGuess.p, (const double*)rowSig, Sold, S, specStart, specEnd, specOk);
    hipLaunchKernelGGL(k_scan_chain, dim3((unsigned)C), dim3(256), 0, st, D, ne, nseg, (const DragSeg*)s->dragSeg.p,
                       start, spec ? (const double*)specStart : nullptr, (const double*)specEnd,
                       (const int*)specOk, S, (double*)s->dragSegStart.p, (int*)s->dragSegMode.p, end, b.slow);
    if (S)
        hipLaunchKernelGGL(k_scan_fill, dim3(wgrid), dim3(256), 0, st, D, ne, C, nseg, (const DragSeg*)s->dragSeg.p,
                           (const double*)s->dragSegStart.p, (const int*)s->dragSegMode.p, start, Sold, S,
                           (int*)b.changed);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

// One frame-dragging pass over an ordered list of bodies (dragOuter, or the
// near part of dragFast): the fixed point described in gx_drag.hip, or one
// sequential thread when it has not settled within dragIters iterations (or
// the chains would not fit).  *b.on must hold spinning() for the pass.
int drag_pass(gx_solver* s, const DragArgs& a, const int32_t* list, long long ne, double* vx, double* vy, double* vz,
              double* jx, double* jy, double* jz, const DragBufs& b, bool holesToo = false) {
    if (a.nh <= 0) return GX_OK;
    hipStream_t st = s->stream;
    const long long C = 3LL * a.nh;
    const long long nseg = (ne + kDragSeg - 1) / kDragSeg;
    const size_t chain = (size_t)C * (size_t)ne * 8;
    int rc;
    if (ne <= kDragSmall && C <= 1024) {  // one workgroup, no host synchronisation
        if (ne > 0) s->dragStats.passes++;
        if ((rc = ensure(s->dragD, std::max<size_t>(chain, 8))) || (rc = ensure(s->dragS0, std::max<size_t>(chain, 8))) ||
            (rc = ensure(s->dragS1, std::max<size_t>(chain, 8))) || (rc = ensure(s->dragSpec, (size_t)C * 16)))
            return rc;
        hipLaunchKernelGGL(k_drag_small, dim3(1), dim3(1024), 0, st, a, list, ne, vx, vy, vz, jx, jy, jz,
                           (const int*)b.on, (double*)s->dragD.p, (double*)s->dragS0.p, (double*)s->dragS1.p,
                           (double*)s->dragSpec.p, s->dragIters, holesToo ? 1 : 0,
                           (unsigned long long*)s->dragCount.p);
        HIPCHK(hipGetLastError());
        return GX_OK;
    }
    if (ne <= 0) return GX_OK;
    s->dragStats.passes++;
    if (s->dragIters > 0 && chain <= ((size_t)8 << 30) && C * nseg < (1LL << 30)) {
        if ((rc = ensure(s->dragD, chain)) || (rc = ensure(s->dragS0, chain)) || (rc = ensure(s->dragS1, chain)) ||
            (rc = ensure(s->dragSig, (size_t)C * nseg * 8 * (1 + kDragRows))) ||
            (rc = ensure(s->dragGuess, (size_t)C * nseg * 8)) ||
            (rc = ensure(s->dragSeg, (size_t)C * nseg * sizeof(DragSeg))) ||
            (rc = ensure(s->dragSegStart, (size_t)C * nseg * 8)) || (rc = ensure(s->dragSegMode, (size_t)C * nseg * 4)) ||
            (rc = ensure(s->dragSpec, (size_t)C * nseg * 20)))
            return rc;
        const unsigned grid = (unsigned)std::min<long long>(blocks(ne, 256), 4096);
        double* D = (double*)s->dragD.p;
        hipLaunchKernelGGL(k_drag_chain_io, dim3(1), dim3(64), 0, st, a.holes, a.nh, jx, jy, jz, b.startJ, 0);
        const double* Sold = nullptr;  // the guess: null is the start spins
        double* Snew = (double*)s->dragS0.p;
        bool settled = false;
        int it = 0;
        // the first kDragApprox iterations only improve the guess (approximate
        // scans); the exact ones check for the fixed point
        constexpr int kDragApprox = 2;
        while (it < s->dragIters && !settled) {
            it++;
            const bool exact = it > kDragApprox;
            if (exact) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, b.changed, 2, 0ULL);  // changed, slow
            hipLaunchKernelGGL(k_drag_eval, dim3(grid), dim3(256), 0, st, a, list, ne, vx, vy, vz, (const double*)jx,
                               (const double*)jy, (const double*)jz, Sold, 0, (const int*)b.on, D);
            if ((rc = drag_scan(s, D, ne, (int)C, b.startJ, Sold, Snew, b.endJ, b, exact))) return rc;
            if (exact) {
                HIPCHK(hipMemcpyAsync(&s->hhist[8], b.changed, 16, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                settled = s->hhist[8] == 0;
                s->dragStats.segments += C * nseg;
                s->dragStats.slow_segments += (int64_t)s->hhist[9];
            }
            if (!settled) {
                Sold = Snew;
                Snew = Snew == (double*)s->dragS0.p ? (double*)s->dragS1.p : (double*)s->dragS0.p;
            }
        }
        s->dragStats.iterations += it;
        s->dragStats.max_iterations = std::max<int64_t>(s->dragStats.max_iterations, it);
        if (settled) {
            // Sold is the reference's chain: the bodies' velocities and the
            // holes' velocity increments from it, then the holes' velocities
            // and spins (the last scan's chain ends).
            hipLaunchKernelGGL(k_drag_eval, dim3(grid), dim3(256), 0, st, a, list, ne, vx, vy, vz, (const double*)jx,
                               (const double*)jy, (const double*)jz, Sold, 1, (const int*)b.on, D);
            hipLaunchKernelGGL(k_drag_chain_io, dim3(1), dim3(64), 0, st, a.holes, a.nh, vx, vy, vz, b.startV, 0);
            if ((rc = drag_scan(s, D, ne, (int)C, b.startV, nullptr, nullptr, b.endV, b, true))) return rc;
            hipLaunchKernelGGL(k_drag_chain_io, dim3(1), dim3(64), 0, st, a.holes, a.nh, vx, vy, vz, b.endV, 1);
            hipLaunchKernelGGL(k_drag_chain_io, dim3(1), dim3(64), 0, st, a.holes, a.nh, jx, jy, jz, b.endJ, 1);
            if (holesToo)
                hipLaunchKernelGGL(k_drag_seq, dim3(1), dim3(64), 0, st, a, a.holes, (long long)a.nh, vx, vy, vz, jx,
                                   jy, jz, (const int*)b.on);
            HIPCHK(hipGetLastError());
            return GX_OK;
        }
    }
    s->dragStats.sequential_passes++;
    hipLaunchKernelGGL(k_drag_seq, dim3(1), dim3(64), 0, st, a, list, ne, vx, vy, vz, jx, jy, jz, (const int*)b.on);
    if (holesToo)
        hipLaunchKernelGGL(k_drag_seq, dim3(1), dim3(64), 0, st, a, a.holes, (long long)a.nh, vx, vy, vz, jx, jy, jz,
                           (const int*)b.on);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

// dragFast (framedrag.go:129-143): the near list as a pass, then the holes one
// by one (they read each other's velocities and spins).
int drag_fast(gx_solver* s, const DragArgs& a, const int32_t* nearList, long long nn, double* vx, double* vy,
              double* vz, double* jx, double* jy, double* jz, const DragBufs& b, const Meta* meta) {
    hipLaunchKernelGGL(k_drag_spinning, dim3(1), dim3(64), 0, s->stream, a.holes, meta, (const double*)jx,
                       (const double*)jy, (const double*)jz, a.c, b.on);
    return drag_pass(s, a, nearList, nn, vx, vy, vz, jx, jy, jz, b, true);
}

// The bodies of dragOuter (alive, not holes, not near; state order) into
// dragList; returns their count (synchronises).
int drag_outer_list(gx_solver* s, long long n, const uint8_t* kind, const uint8_t* alive, const uint8_t* nearFlag,
                    long long* count) {
    const long long nb = std::max(1LL, (n + kCollectTile - 1) / kCollectTile);
    int rc;
    if ((rc = ensure(s->dragFlag, (size_t)std::max(n, 1LL))) || (rc = ensure(s->dragList, (size_t)std::max(n, 1LL) * 4)) ||
        (rc = ensure(s->tileCnt, (size_t)nb * 4)) || (rc = ensure(s->splitScalars, 32)))
        return rc;
    hipStream_t st = s->stream;
    const unsigned grid = (unsigned)std::min<long long>(blocks(n, 256), 4096);
    long long* cnt = (long long*)s->splitScalars.p + 2;
    hipLaunchKernelGGL(k_drag_outer_flag, dim3(grid), dim3(256), 0, st, n, kind, alive, nearFlag,
                       (uint8_t*)s->dragFlag.p);
    hipLaunchKernelGGL(k_flag_count, dim3((unsigned)nb), dim3(256), 0, st, n, (const uint8_t*)s->dragFlag.p,
                       (int*)s->tileCnt.p);
    hipLaunchKernelGGL(k_flag_scan, dim3(1), dim3(1024), 0, st, (int)nb, (int*)s->tileCnt.p, cnt);
    hipLaunchKernelGGL(k_flag_scatter, dim3((unsigned)nb), dim3(256), 0, st, n, (const uint8_t*)s->dragFlag.p,
                       (const int*)s->tileCnt.p, (int32_t*)s->dragList.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&s->hhist[10], cnt, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *count = (long long)s->hhist[10];
    return GX_OK;
}

// dragOuter (framedrag.go:115-127) over the list built by drag_outer_list.
int drag_outer(gx_solver* s, const DragArgs& a, long long count, double* vx, double* vy, double* vz, double* jx,
               double* jy, double* jz, const DragBufs& b, const Meta* meta) {
    hipLaunchKernelGGL(k_drag_spinning, dim3(1), dim3(64), 0, s->stream, a.holes, meta, (const double*)jx,
                       (const double*)jy, (const double*)jz, a.c, b.on);
    return drag_pass(s, a, (const int32_t*)s->dragList.p, count, vx, vy, vz, jx, jy, jz, b);
}

// holeHistory.push, retarded.go:75-83, the header part: the ring slot that the
// new entry goes to.
int hist_advance(gx_retard::Hist& h, double t) {
    h.head = (h.head + 1) % kHistLen;
    if (h.n < kHistLen) h.n++;
    h.t0 = t - (double)(h.n - 1) * h.dt;
    return h.head;
}

gx_retard::Hist* hist_find(gx_retard* r, uint64_t id) {
    for (auto& h : r->hists)
        if (h.id == id) return &h;
    return nullptr;
}

// history, retarded.go:163-179: a new, empty history (the caller pushes the
// current state as its first entry).
int hist_create(gx_retard* r, uint64_t id, double dt, gx_retard::Hist** out) {
    gx_retard::Hist h{id, nullptr, 0.0, -INFINITY, dt, 0, 0};
    HIPCHK(hipMalloc(&h.ring, (size_t)6 * kHistLen * sizeof(double)));
    r->hists.push_back(h);
    *out = &r->hists.back();
    return GX_OK;
}

void hists_clear(gx_retard* r) {
    for (auto& h : r->hists) hipFree(h.ring);
    r->hists.clear();
}

// Writes the listed pushes into the rings (k_retard_push) and waits.
int run_pushes(gx_solver* s, gx_retard* r, const std::vector<HistPush>& p, const double* x, const double* y,
               const double* z, const double* vx, const double* vy, const double* vz, double c) {
    if (p.empty()) return GX_OK;
    int rc;
    if ((rc = ensure(r->pushes, p.size() * sizeof(HistPush)))) return rc;
    HIPCHK(hipMemcpyAsync(r->pushes.p, p.data(), p.size() * sizeof(HistPush), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_retard_push, dim3((unsigned)std::min<size_t>((p.size() + 63) / 64, 1024)), dim3(64), 0,
                       s->stream, (const HistPush*)r->pushes.p, (int)p.size(), x, y, z, vx, vy, vz, c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

// The alive holes of the last collect (slot order) and their IDs, to the host.
int fetch_holes(gx_solver* s, gx_retard* r, long long nh, const uint64_t* ids) {
    r->holeIdx.resize((size_t)nh);
    r->holeIds.resize((size_t)nh);
    if (nh == 0) return GX_OK;
    int rc;
    if ((rc = ensure(r->ids, (size_t)nh * 8))) return rc;
    hipLaunchKernelGGL(k_gather_hole_ids, dim3((unsigned)std::min<long long>(blocks(nh, 256), 1024)), dim3(256), 0,
                       s->stream, (const int32_t*)s->holes.p, nh, ids, (uint64_t*)r->ids.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(r->holeIdx.data(), s->holes.p, (size_t)nh * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(r->holeIds.data(), r->ids.p, (size_t)nh * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

bool retarded(const gx_retard* r, int mode, double c) { return r && mode != GX_GRAVITY_INSTANT && c > 0; }

// bindHistories, retarded.go:106-116: the history of every hole slot (after
// fetch_holes) for the kernels.  Instant gravity: rd.hist stays null.
int bind_histories(gx_solver* s, gx_retard* r, long long nh, int mode, double time, double dt, double c,
                   RetardDev* rd) {
    rd->hist = nullptr;
    rd->mode = mode;
    rd->time = time;
    rd->dt = dt;
    if (!retarded(r, mode, c) || nh == 0) return GX_OK;
    std::vector<HistDev> hd((size_t)nh);
    for (long long k = 0; k < nh; k++) {
        const gx_retard::Hist* h = hist_find(r, r->holeIds[(size_t)k]);
        hd[(size_t)k] = h ? HistDev{h->ring, h->t0, h->appeared, h->dt, h->n, h->head}
                          : HistDev{nullptr, 0.0, 0.0, 0.0, 0, 0};
    }
    int rc;
    if ((rc = ensure(r->bound, hd.size() * sizeof(HistDev)))) return rc;
    HIPCHK(hipMemcpyAsync(r->bound.p, hd.data(), hd.size() * sizeof(HistDev), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    rd->hist = (const HistDev*)r->bound.p;
    return GX_OK;
}

// recordHoles, retarded.go:181-209, at the time after the step: every alive
// hole (the slots of fetch_holes) appends its state to its history, a hole
// without one starts it; histories of holes gone are dropped.  Outside the
// retarded modes the histories are cleared.
int record_holes(gx_solver* s, gx_retard* r, const double* x, const double* y, const double* z, const double* vx,
                 const double* vy, const double* vz, double time, double dt, double c, int mode) {
    if (!r) return GX_OK;
    if (!retarded(r, mode, c)) {
        hists_clear(r);
        return GX_OK;
    }
    const size_t nh = r->holeIds.size();
    const bool seen = r->hists.size() != nh;
    std::vector<HistPush> p;
    int rc;
    for (size_t k = 0; k < nh; k++) {
        gx_retard::Hist* h = hist_find(r, r->holeIds[k]);
        if (!h && (rc = hist_create(r, r->holeIds[k], dt, &h))) return rc;
        p.push_back(HistPush{h->ring, r->holeIdx[k], hist_advance(*h, time)});
    }
    if ((rc = run_pushes(s, r, p, x, y, z, vx, vy, vz, c))) return rc;
    if (seen) {  // drop histories of holes that merged away
        size_t keep = 0;
        for (size_t j = 0; j < r->hists.size(); j++) {
            const bool alive = std::find(r->holeIds.begin(), r->holeIds.end(), r->hists[j].id) != r->holeIds.end();
            if (alive)
                r->hists[keep++] = r->hists[j];
            else
                hipFree(r->hists[j].ring);
        }
        r->hists.resize(keep);
    }
    return GX_OK;
}

// One step with black holes, stepSplit (internal/sim/split.go:164-199) for a
// state without gas; without alive holes it is step()'s plain KDK
// (sim.go:266-277).  With a history set r in a retarded mode the holes pull
// from their sources on the light cone (retarded.go), and the holes are
// recorded at time + dt after the step (sim.go:285-287).  With c > 0 the pairs with holes use
// Paczynski-Wiita radii and drifts use the proper velocity, and with spins
// (jx non-null) the holes drag frames (framedrag.go) around the outer kicks
// and the substeps' drifts.
int split_step(gx_solver* s, StepArgs a, double* vx, double* vy, double* vz, double* jx, double* jy, double* jz,
               double dt, double c, double eps, int accValid, int32_t* substepsOut, gx_retard* retard = nullptr,
               const uint64_t* ids = nullptr, double time = 0, int mode = GX_GRAVITY_INSTANT) {
    const long long n = a.n;
    int sub = 1;
    if (substepsOut) *substepsOut = 1;
    if (n == 0) {
        if (retard) {
            retard->holeIds.clear();
            retard->holeIdx.clear();
        }
        return record_holes(s, retard, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, time + dt, dt, c, mode);
    }
    double* x = (double*)a.x;
    double* y = (double*)a.y;
    double* z = (double*)a.z;
    const long long nb = std::max(1LL, (n + kCollectTile - 1) / kCollectTile);
    int rc;
    if ((rc = reserve(s, n)) || (rc = ensure(s->nearFlag, (size_t)n)) || (rc = ensure(s->nearList, (size_t)n * 4)) ||
        (rc = ensure(s->tileCnt, (size_t)nb * 4)) || (rc = ensure(s->splitScalars, 32)) ||
        (rc = ensure(s->farF, (size_t)n * 24)) || (rc = ensure(s->fastF, (size_t)n * 24)))
        return rc;
    // slices of the holes' pair sums: ~4096 bodies per workgroup over the far bodies
    const int farSlices = (int)std::min<long long>(256, std::max(1LL, n / 4096));
    if (!accValid && (rc = run_dev(s, a))) return rc;  // initForces
    hipStream_t st = s->stream;
    const unsigned grid = (unsigned)std::min<long long>(blocks(n, 256), 4096);
    const double eps2 = eps * eps;
    // the alive holes in state order (collectHoles, sim.go:291-299)
    enqueue_collect(s, a, false);
    HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const long long nh = s->hmeta->nholes;
    if ((rc = ensure(s->partial, (size_t)std::max<long long>(nh, 1) * farSlices * 24))) return rc;
    double* partial = (double*)s->partial.p;
    RetardDev rd{nullptr, mode, time, dt};
    if (retard && ((rc = fetch_holes(s, retard, retarded(retard, mode, c) ? nh : 0, ids)) ||
                   (rc = bind_histories(s, retard, nh, mode, time, dt, c, &rd))))
        return rc;
    if (nh == 0) {
        // kick, drift (sim.go:316-343: dt / sqrt(1 + u^2 / c^2) with c > 0), forces, kick
        if ((rc = kick(s, vx, vy, vz, a.ax, a.ay, a.az, n, 0.5 * dt))) return rc;
        hipLaunchKernelGGL(k_split_drift_far, dim3(grid), dim3(256), 0, st, n, a.kind, (const uint8_t*)nullptr, x, y, z,
                           (const double*)vx, (const double*)vy, (const double*)vz, dt, c);
        HIPCHK(hipGetLastError());
        if ((rc = run_dev(s, a)) || (rc = kick(s, vx, vy, vz, a.ax, a.ay, a.az, n, 0.5 * dt))) return rc;
        HIPCHK(hipStreamSynchronize(st));
        return record_holes(s, retard, x, y, z, vx, vy, vz, time + dt, dt, c, mode);
    }
    const Meta* meta = (const Meta*)s->meta.p;
    const int32_t* holes = (const int32_t*)s->holes.p;
    uint8_t* nearFlag = (uint8_t*)s->nearFlag.p;
    int32_t* nearList = (int32_t*)s->nearList.p;
    long long* nnDev = (long long*)s->splitScalars.p;
    unsigned long long* tauEnc = (unsigned long long*)s->splitScalars.p + 1;
    double* farX = (double*)s->farF.p;
    double* farY = farX + n;
    double* farZ = farY + n;
    double* fastX = (double*)s->fastF.p;
    double* fastY = fastX + n;
    double* fastZ = fastY + n;
    // classify (split.go:39-80)
    const double inf = INFINITY;
    unsigned long long infBits;
    memcpy(&infBits, &inf, 8);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, tauEnc, 1, infBits | 0x8000000000000000ull);
    hipLaunchKernelGGL(k_split_classify, dim3(grid), dim3(256), 0, st, n, (const double*)x, (const double*)y,
                       (const double*)z, a.m, a.kind, holes, meta, eps2, c, 20.0 * dt, nearFlag, tauEnc);
    hipLaunchKernelGGL(k_flag_count, dim3((unsigned)nb), dim3(256), 0, st, n, (const uint8_t*)nearFlag,
                       (int*)s->tileCnt.p);
    hipLaunchKernelGGL(k_flag_scan, dim3(1), dim3(1024), 0, st, (int)nb, (int*)s->tileCnt.p, nnDev);
    hipLaunchKernelGGL(k_flag_scatter, dim3((unsigned)nb), dim3(256), 0, st, n, (const uint8_t*)nearFlag,
                       (const int*)s->tileCnt.p, nearList);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s->hhist, s->splitScalars.p, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const unsigned long long te = s->hhist[1];
    const unsigned long long tb = (te >> 63) ? (te & 0x7fffffffffffffffull) : ~te;
    double tauMin;
    memcpy(&tauMin, &tb, 8);
    if (!std::isinf(tauMin)) {
        while ((double)sub * tauMin < 20.0 * dt && sub < 256) sub *= 2;
    }
    if (substepsOut) *substepsOut = sub;
    // frame dragging (framedrag.go): acts only with c > 0 and a spinning hole
    const bool dragOn = jx && c > 0;
    const long long nnHost = (long long)s->hhist[0];
    DragArgs da{x, y, z, a.m, a.alive, holes, (int)nh, eps2, c, 0.0};
    DragBufs db{};
    long long nOuter = 0;
    if (dragOn && ((rc = drag_bufs(s, nh, db)) || (rc = drag_outer_list(s, n, a.kind, a.alive, nearFlag, &nOuter))))
        return rc;
    auto drag_o = [&](double h) -> int {  // dragOuter, framedrag.go:115-127
        if (!dragOn) return GX_OK;
        DragArgs d = da;
        d.dt = h;
        return drag_outer(s, d, nOuter, vx, vy, vz, jx, jy, jz, db, meta);
    };
    auto drag_f = [&](double h) -> int {  // dragFast, framedrag.go:129-143
        if (!dragOn) return GX_OK;
        DragArgs d = da;
        d.dt = h;
        return drag_fast(s, d, (const int32_t*)nearList, nnHost, vx, vy, vz, jx, jy, jz, db, meta);
    };
    auto outer = [&]() {  // outerHoleForces, split.go:82-117
        hipLaunchKernelGGL(k_split_far, dim3(grid), dim3(256), 0, st, n, (const double*)x, (const double*)y,
                           (const double*)z, a.m, a.kind, (const uint8_t*)nearFlag, holes, meta, eps2, c, rd, farX,
                           farY, farZ);
        hipLaunchKernelGGL(k_split_hole_sum, dim3((unsigned)nh, (unsigned)farSlices), dim3(256), 0, st, n,
                           (const double*)x, (const double*)y, (const double*)z, a.m, a.kind, a.dmOff,
                           (const uint8_t*)nearFlag, (const int32_t*)nearList, (const long long*)nnDev, holes, meta,
                           eps2, c, 0, partial);
        hipLaunchKernelGGL(k_split_hole_finish, dim3(1), dim3(64), 0, st, (const double*)x, (const double*)y,
                           (const double*)z, a.m, a.kind, holes, meta, eps2, c, rd, 0, farSlices,
                           (const double*)partial, farX, farY, farZ);
    };
    auto fast = [&]() {  // fastForces, split.go:141-160
        hipLaunchKernelGGL(k_split_fast, dim3(grid), dim3(256), 0, st, (const int32_t*)nearList,
                           (const long long*)nnDev, (const double*)x, (const double*)y, (const double*)z, a.m, a.kind,
                           holes, meta, eps2, c, rd, fastX, fastY, fastZ);
        hipLaunchKernelGGL(k_split_hole_sum, dim3((unsigned)nh, 1), dim3(256), 0, st, n, (const double*)x,
                           (const double*)y, (const double*)z, a.m, a.kind, a.dmOff, (const uint8_t*)nearFlag,
                           (const int32_t*)nearList, (const long long*)nnDev, holes, meta, eps2, c, 1, partial);
        hipLaunchKernelGGL(k_split_hole_finish, dim3(1), dim3(64), 0, st, (const double*)x, (const double*)y,
                           (const double*)z, a.m, a.kind, holes, meta, eps2, c, rd, 1, 1, (const double*)partial,
                           fastX, fastY, fastZ);
    };
    auto kick_outer = [&](double h) {  // kickOuter, split.go:201-208
        hipLaunchKernelGGL(k_split_kick_outer, dim3(grid), dim3(256), 0, st, n, vx, vy, vz, (const double*)a.ax,
                           (const double*)a.ay, (const double*)a.az, (const double*)farX, (const double*)farY,
                           (const double*)farZ, h);
    };
    auto kick_fast = [&](double h) {  // kickFast, split.go:210-219: near stars, then holes
        hipLaunchKernelGGL(k_split_kick_list, dim3(grid), dim3(256), 0, st, (const int32_t*)nearList, 0LL,
                           (const long long*)nnDev, vx, vy, vz, (const double*)fastX, (const double*)fastY,
                           (const double*)fastZ, h);
        hipLaunchKernelGGL(k_split_kick_list, dim3(1), dim3(256), 0, st, holes, nh, (const long long*)nullptr, vx, vy,
                           vz, (const double*)fastX, (const double*)fastY, (const double*)fastZ, h);
    };
    outer();
    kick_outer(0.5 * dt);
    if ((rc = drag_o(0.5 * dt))) return rc;
    const double h = dt / (double)sub;
    fast();
    for (int k = 0; k < sub; k++) {
        kick_fast(0.5 * h);
        if ((rc = drag_f(0.5 * h))) return rc;
        hipLaunchKernelGGL(k_split_drift_list, dim3(grid), dim3(256), 0, st, (const int32_t*)nearList, 0LL,
                           (const long long*)nnDev, x, y, z, (const double*)vx, (const double*)vy, (const double*)vz,
                           h, c);
        hipLaunchKernelGGL(k_split_drift_list, dim3(1), dim3(256), 0, st, holes, nh, (const long long*)nullptr, x, y, z,
                           (const double*)vx, (const double*)vy, (const double*)vz, h, c);
        fast();
        if ((rc = drag_f(0.5 * h))) return rc;
        kick_fast(0.5 * h);
    }
    hipLaunchKernelGGL(k_split_drift_far, dim3(grid), dim3(256), 0, st, n, a.kind, (const uint8_t*)nearFlag, x, y, z,
                       (const double*)vx, (const double*)vy, (const double*)vz, dt, c);
    HIPCHK(hipGetLastError());
    if ((rc = run_dev(s, a))) return rc;  // the slow forces at the new positions
    outer();
    if ((rc = drag_o(0.5 * dt))) return rc;
    kick_outer(0.5 * dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return record_holes(s, retard, x, y, z, vx, vy, vz, time + dt, dt, c, mode);  // sim.go:285-287
}

}  // namespace

extern "C" {

int gx_split_step_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                      const double* mass, const uint8_t* kind, const uint8_t* alive, int64_t n, double eps,
                      int dark_matter_off, double dt, double c, int acc_valid, double* ax, double* ay, double* az,
                      int32_t* substeps) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) || (n > 0 && (!vx || !vy || !vz) && (rc = set_err(GX_ERR_ARG, "null array"))) ||
        (rc = set_single(s, n, s->theta, eps)))
        return rc;
    return split_step(s, a, vx, vy, vz, nullptr, nullptr, nullptr, dt, c, eps, acc_valid, substeps);
}

int gx_split_step_spin_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                           const double* mass, const uint8_t* kind, const uint8_t* alive, double* spin_x,
                           double* spin_y, double* spin_z, int64_t n, double eps, int dark_matter_off, double dt,
                           double c, int acc_valid, double* ax, double* ay, double* az, int32_t* substeps) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) ||
        (n > 0 && (!vx || !vy || !vz || !spin_x || !spin_y || !spin_z) && (rc = set_err(GX_ERR_ARG, "null array"))) ||
        (rc = set_single(s, n, s->theta, eps)))
        return rc;
    return split_step(s, a, vx, vy, vz, n > 0 ? spin_x : nullptr, spin_y, spin_z, dt, c, eps, acc_valid, substeps);
}

int gx_split_step_retarded_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                               const double* mass, const uint8_t* kind, const uint8_t* alive, double* spin_x,
                               double* spin_y, double* spin_z, const uint64_t* ids, int64_t n, double eps,
                               int dark_matter_off, double dt, double c, int acc_valid, double* ax, double* ay,
                               double* az, gx_retard* r, double time, int gravity, int32_t* substeps) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (gravity < GX_GRAVITY_INSTANT || gravity > GX_GRAVITY_RETARDED_V)
        return set_err(GX_ERR_ARG, "unknown gravity mode %d (want instant, retarded, retarded+v)", gravity);
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) || (n > 0 && (!vx || !vy || !vz) && (rc = set_err(GX_ERR_ARG, "null array"))) ||
        (n > 0 && spin_x && (!spin_y || !spin_z) && (rc = set_err(GX_ERR_ARG, "null spin array"))) ||
        (rc = set_single(s, n, s->theta, eps)))
        return rc;
    return split_step(s, a, vx, vy, vz, n > 0 ? spin_x : nullptr, spin_y, spin_z, dt, c, eps, acc_valid, substeps, r,
                      ids, time, gravity);
}

gx_retard* gx_retard_new(void) { return new (std::nothrow) gx_retard(); }

void gx_retard_free(gx_retard* r) {
    if (!r) return;
    hists_clear(r);
    for (DevBuf* b : {&r->bound, &r->pushes, &r->ids})
        if (b->p) hipFree(b->p);
    delete r;
}

int gx_retard_count(const gx_retard* r) { return r ? (int)r->hists.size() : 0; }

int gx_retard_hole_appeared_dev(gx_solver* s, gx_retard* r, const double* x, const double* y, const double* z,
                                const double* vx, const double* vy, const double* vz, const uint64_t* ids, int64_t n,
                                int64_t i, double time, double dt, double c) {
    if (!s || !r) return set_err(GX_ERR_ARG, "null solver or history set");
    if (i < 0 || i >= n) return set_err(GX_ERR_ARG, "body %lld out of range [0, %lld)", (long long)i, (long long)n);
    if (!x || !y || !z || !vx || !vy || !vz) return set_err(GX_ERR_ARG, "null array");
    uint64_t id = (uint64_t)i;
    if (ids) {
        HIPCHK(hipMemcpyAsync(&id, ids + i, 8, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    gx_retard::Hist* h = hist_find(r, id);
    int rc;
    if (!h) {
        if ((rc = hist_create(r, id, dt, &h))) return rc;
        const std::vector<HistPush> p{HistPush{h->ring, (int32_t)i, hist_advance(*h, time)}};
        if ((rc = run_pushes(s, r, p, x, y, z, vx, vy, vz, c))) return rc;
    }
    h->appeared = time;
    return GX_OK;
}

int gx_retard_source_pos_dev(gx_solver* s, gx_retard* r, const double* x, const double* y, const double* z,
                             const uint8_t* kind, const uint8_t* alive, const uint64_t* ids, int64_t n, int64_t k,
                             const double* obs_x, const double* obs_y, const double* obs_z, int64_t m, double time,
                             double dt, double c, int gravity, double* src_x, double* src_y, double* src_z,
                             uint8_t* visible) {
    if (!s || !r) return set_err(GX_ERR_ARG, "null solver or history set");
    if (gravity < GX_GRAVITY_INSTANT || gravity > GX_GRAVITY_RETARDED_V)
        return set_err(GX_ERR_ARG, "unknown gravity mode %d", gravity);
    if (n <= 0 || !x || !y || !z || !kind || !alive) return set_err(GX_ERR_ARG, "no bodies");
    if (m < 0 || (m > 0 && (!obs_x || !obs_y || !obs_z || !src_x || !src_y || !src_z || !visible)))
        return set_err(GX_ERR_ARG, "null observer array");
    StepArgs a = step_args(x, y, z, nullptr, kind, alive, n, 0, nullptr, nullptr, nullptr);
    int rc;
    if ((rc = set_single(s, n, s->theta, 0.0)) || (rc = reserve(s, n))) return rc;
    enqueue_collect(s, a, false);
    HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const long long nh = s->hmeta->nholes;
    if (k < 0 || k >= nh) return set_err(GX_ERR_ARG, "hole %lld out of range [0, %lld)", (long long)k, nh);
    RetardDev rd{nullptr, gravity, time, dt};
    if ((rc = fetch_holes(s, r, nh, ids)) || (rc = bind_histories(s, r, nh, gravity, time, dt, c, &rd))) return rc;
    if (m == 0) return GX_OK;
    hipLaunchKernelGGL(k_retard_source, dim3((unsigned)std::min<long long>(blocks(m, 256), 4096)), dim3(256), 0,
                       s->stream, rd, (long long)k, (const int32_t*)s->holes.p, x, y, z, obs_x, obs_y, obs_z,
                       (long long)m, c, src_x, src_y, src_z, visible);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

// Shared set-up of the frame-dragging entry points: the alive holes of the
// state (collect), their count, the chain scalars.
static int drag_setup(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                      const uint8_t* kind, const uint8_t* alive, double* vx, double* vy, double* vz, double* jx,
                      double* jy, double* jz, int64_t n, double eps, DragBufs& b, long long* nh) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (n > 0 && (!x || !y || !z || !mass || !kind || !alive || !vx || !vy || !vz || !jx || !jy || !jz))
        return set_err(GX_ERR_ARG, "null array");
    if (n > ((1LL << 31) - 1)) return set_err(GX_ERR_ARG, "n must be below 2^31");
    *nh = 0;
    if (n == 0) return GX_OK;
    int rc;
    if ((rc = set_single(s, n, s->theta, eps)) || (rc = reserve(s, n))) return rc;
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, 0, nullptr, nullptr, nullptr);
    enqueue_collect(s, a, false);
    HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *nh = s->hmeta->nholes;
    return drag_bufs(s, *nh, b);
}

int gx_drag_outer_dev(gx_solver* s, const double* x, const double* y, const double* z, double* vx, double* vy,
                      double* vz, const double* mass, const uint8_t* kind, const uint8_t* alive, const uint8_t* near,
                      double* spin_x, double* spin_y, double* spin_z, int64_t n, double eps, double dt, double c) {
    DragBufs b{};
    long long nh = 0;
    int rc = drag_setup(s, x, y, z, mass, kind, alive, vx, vy, vz, spin_x, spin_y, spin_z, n, eps, b, &nh);
    if (rc || nh == 0) return rc;
    long long count = 0;
    if ((rc = drag_outer_list(s, n, kind, alive, near, &count))) return rc;
    const DragArgs a{x, y, z, mass, alive, (const int32_t*)s->holes.p, (int)nh, eps * eps, c, dt};
    if ((rc = drag_outer(s, a, count, vx, vy, vz, spin_x, spin_y, spin_z, b, (const Meta*)s->meta.p))) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

int gx_drag_fast_dev(gx_solver* s, const double* x, const double* y, const double* z, double* vx, double* vy,
                     double* vz, const double* mass, const uint8_t* kind, const uint8_t* alive,
                     const int32_t* near_list, int64_t near_count, double* spin_x, double* spin_y, double* spin_z,
                     int64_t n, double eps, double dt, double c) {
    if (near_count < 0 || (near_count > 0 && !near_list)) return set_err(GX_ERR_ARG, "bad near list");
    DragBufs b{};
    long long nh = 0;
    int rc = drag_setup(s, x, y, z, mass, kind, alive, vx, vy, vz, spin_x, spin_y, spin_z, n, eps, b, &nh);
    if (rc || nh == 0) return rc;
    const DragArgs a{x, y, z, mass, alive, (const int32_t*)s->holes.p, (int)nh, eps * eps, c, dt};
    if ((rc = drag_fast(s, a, near_list, near_count, vx, vy, vz, spin_x, spin_y, spin_z, b, (const Meta*)s->meta.p)))
        return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

int gx_drag_get_stats(const gx_solver* s, gx_drag_stats* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    *out = s->dragStats;
    if (s->dragCount.p) {  // the one-workgroup passes count on the device
        unsigned long long c[2];
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMemcpy(c, s->dragCount.p, 16, hipMemcpyDeviceToHost));
        out->iterations += (int64_t)c[0];
        out->sequential_passes += (int64_t)c[1];
    }
    return GX_OK;
}

int gx_accelerations_active_dev(gx_solver* s, const double* x, const double* y, const double* z,
                                const double* mass, const uint8_t* kind, const uint8_t* alive,
                                const uint8_t* active, int64_t n, double eps, int dark_matter_off, double* ax,
                                double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (n > 0 && !active) return set_err(GX_ERR_ARG, "null active mask");
    StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    a.active = active;
    int rc;
    if ((rc = check_arrays(a)) || (rc = set_single(s, n, s->theta, eps))) return rc;
    return run_dev(s, a);
}

int gx_block_step_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                      const double* mass, const uint8_t* kind, const uint8_t* alive, int64_t n, double eps,
                      int dark_matter_off, double T, int32_t levels, double eta, int acc_valid, double* ax,
                      double* ay, double* az, gx_block_stats* out) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) || (n > 0 && (!vx || !vy || !vz) && (rc = set_err(GX_ERR_ARG, "null array"))) ||
        (rc = set_single(s, n, s->theta, eps)))
        return rc;
    return block_step(s, a, vx, vy, vz, T, levels, eta, eps, acc_valid, out);
}

int gx_diagnostics_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* vx,
                       const double* vy, const double* vz, const double* mass, const uint8_t* kind, int64_t n,
                       int dark_matter_off, double* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    if (n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !mass || !kind)) return set_err(GX_ERR_ARG, "null array");
    if (s->pending) (void)finish_step(s);
    for (int c = 0; c < GX_DIAG_VALUES; c++) out[c] = 0;
    if (n == 0) return GX_OK;
    const int nb = (int)std::min<long long>(blocks(n, 256), kDiagBlocks);
    int rc;
    if ((rc = ensure(s->diag, (size_t)nb * kDiagValues * 8))) return rc;
    hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, s->stream, (long long)n, x, y, z, vx, vy, vz, mass, kind,
                       dark_matter_off, (double*)s->diag.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h((size_t)nb * kDiagValues);
    HIPCHK(hipMemcpyAsync(h.data(), s->diag.p, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double acc[kDiagValues] = {};
    for (int b = 0; b < nb; b++)  // fixed order: the result does not depend on timing
        for (int c = 0; c < kDiagValues; c++) acc[c] += h[(size_t)b * kDiagValues + c];
    for (int c = 0; c < kDiagValues; c++) out[c] = acc[c];
    return GX_OK;
}

int gx_accelerations_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                         const uint8_t* kind, const uint8_t* alive, int64_t n, double eps, int dark_matter_off,
                         double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) || (rc = set_single(s, n, s->theta, eps))) return rc;
    return run_dev(s, a);
}

int gx_enqueue_accelerations_dev(gx_solver* s, const double* x, const double* y, const double* z,
                                 const double* mass, const uint8_t* kind, const uint8_t* alive, int64_t n,
                                 double eps, int dark_matter_off, double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) || (rc = set_single(s, n, s->theta, eps))) return rc;
    s->pending = false;  // a new step supersedes an unsynchronised one
    s->allLevels = true;  // no repeat is possible once the caller has moved on
    rc = enqueue_step(s, a);
    s->allLevels = false;
    return rc;
}

int gx_sync(gx_solver* s) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (!s->pending) {
        HIPCHK(hipStreamSynchronize(s->stream));
        return GX_OK;
    }
    int rc = finish_step(s);
    if (rc) return rc;
    if (s->stats.flags & (kFlagStackOverflow | kFlagListOverflow))
        return set_err(GX_ERR_CAPACITY, "walk stack or list overflow; use gx_accelerations_dev");
    return GX_OK;
}

int gx_accelerations(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                     const uint8_t* kind, const uint8_t* alive, int64_t n, double eps, int dark_matter_off,
                     double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (n == 0) return GX_OK;
    if (!x || !y || !z || !mass || !kind || !alive || !ax || !ay || !az) return set_err(GX_ERR_ARG, "null array");
    int rc;
    const size_t d = (size_t)n * 8, b = (size_t)n;
    if ((rc = ensure(s->x, d)) || (rc = ensure(s->y, d)) || (rc = ensure(s->z, d)) || (rc = ensure(s->m, d)) ||
        (rc = ensure(s->kind, b)) || (rc = ensure(s->alive, b)) || (rc = ensure(s->ax, d)) ||
        (rc = ensure(s->ay, d)) || (rc = ensure(s->az, d)))
        return rc;
    hipStream_t st = s->stream;
    HIPCHK(hipMemcpyAsync(s->x.p, x, d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->y.p, y, d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->z.p, z, d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->m.p, mass, d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->kind.p, kind, b, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->alive.p, alive, b, hipMemcpyHostToDevice, st));
    // dead bodies keep their values: start from the caller's arrays
    HIPCHK(hipMemcpyAsync(s->ax.p, ax, d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->ay.p, ay, d, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->az.p, az, d, hipMemcpyHostToDevice, st));
    rc = gx_accelerations_dev(s, (const double*)s->x.p, (const double*)s->y.p, (const double*)s->z.p,
                              (const double*)s->m.p, (const uint8_t*)s->kind.p, (const uint8_t*)s->alive.p, n, eps,
                              dark_matter_off, (double*)s->ax.p, (double*)s->ay.p, (double*)s->az.p);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ax, s->ax.p, d, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ay, s->ay.p, d, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(az, s->az.p, d, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return GX_OK;
}

int gx_batch_accelerations_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                               const uint8_t* kind, const uint8_t* alive, const int64_t* offsets, int32_t count,
                               const double* theta, const double* eps, int dark_matter_off, double* ax, double* ay,
                               double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    int rc;
    if ((rc = check_batch(offsets, count, theta, eps))) return rc;
    const StepArgs a = step_args(x, y, z, mass, kind, alive, offsets[count], dark_matter_off, ax, ay, az);
    if ((rc = check_arrays(a)) || (rc = set_instances(s, count, offsets, theta, 0, eps, 0))) return rc;
    return run_dev(s, a);
}

int gx_kick_dev(gx_solver* s, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                const double* az, int64_t n, double h) {
    if (!s || n < 0) return set_err(GX_ERR_ARG, "bad argument");
    return kick(s, vx, vy, vz, ax, ay, az, n, h);
}

int gx_drift_dev(gx_solver* s, double* x, double* y, double* z, const double* vx, const double* vy,
                 const double* vz, int64_t n, double dt) {
    if (!s || n < 0) return set_err(GX_ERR_ARG, "bad argument");
    return drift(s, x, y, z, vx, vy, vz, n, dt);
}

int gx_step_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                const double* mass, const uint8_t* kind, const uint8_t* alive, int64_t n, double eps,
                int dark_matter_off, double dt, int acc_valid, double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    int rc;
    if ((rc = check_arrays(a)) || (rc = set_single(s, n, s->theta, eps))) return rc;
    return kdk(s, a, vx, vy, vz, dt, acc_valid);
}

int gx_batch_step_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                      const double* mass, const uint8_t* kind, const uint8_t* alive, const int64_t* offsets,
                      int32_t count, const double* theta, const double* eps, int dark_matter_off, double dt,
                      int acc_valid, double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    int rc;
    if ((rc = check_batch(offsets, count, theta, eps))) return rc;
    const StepArgs a = step_args(x, y, z, mass, kind, alive, offsets[count], dark_matter_off, ax, ay, az);
    if ((rc = check_arrays(a)) || (rc = set_instances(s, count, offsets, theta, 0, eps, 0))) return rc;
    return kdk(s, a, vx, vy, vz, dt, acc_valid);
}

int gx_solver_autotune(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                       const uint8_t* kind, const uint8_t* alive, const int64_t* offsets, int32_t count,
                       const double* theta, const double* eps, int dark_matter_off, double* ax, double* ay,
                       double* az, int32_t reps, gx_tune_entry* table, int32_t table_len, int32_t* entries) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    int rc;
    if ((rc = check_batch(offsets, count, theta, eps))) return rc;
    if (reps < 1 || reps > 1000 || table_len < 0 || (table_len > 0 && !table))
        return set_err(GX_ERR_ARG, "bad repetitions or table");
    if (entries) *entries = 0;
    const StepArgs a = step_args(x, y, z, mass, kind, alive, offsets[count], dark_matter_off, ax, ay, az);
    if ((rc = check_arrays(a)) || (rc = set_instances(s, count, offsets, theta, 0, eps, 0))) return rc;
    if (s->walk == GX_WALK_FP64 || a.n == 0) return GX_OK;  // nothing to tune
    auto report = [&](const std::vector<gx_tune_entry>& t) {
        const int k = std::min<int>((int)t.size(), table_len);
        for (int i = 0; i < k; i++) table[i] = t[(size_t)i];
        if (entries) *entries = (int32_t)t.size();
    };
    for (const auto& c : s->tuned)  // a workload of this shape was tuned already
        if (c.n == a.n && c.B == count && c.walk == s->walk && c.monopole == s->monopole) {
            s->gsize = c.gsize;
            s->persistent = c.persistent;
            s->walkConfig = c.config;
            report(c.table);
            return GX_OK;
        }
    const int timing = s->timing;
    s->timing = 0;
    std::vector<gx_tune_entry> t;
    for (int gsize : {64, 32}) {
        s->gsize = gsize;
        s->persistent = 1;
        s->walkConfig = kWalkDefault;
        if ((rc = run_dev(s, a))) break;  // the tree and this group size's walk groups
        // the step allocated Meta on a fresh solver: take the pointer only now
        Meta* meta = (Meta*)s->meta.p;
        if (!meta) {
            rc = set_err(GX_ERR_HIP, "no step metadata after a force step");
            break;
        }
        for (int cfg = 0; cfg < kWalkConfigs && !rc; cfg++)
            for (int pers : {0, 2}) {
                s->walkConfig = cfg;
                s->persistent = pers;
                bool done = false;
                for (int attempt = 0; attempt < 4 && !done; attempt++) {
                    // one untimed walk, then reps timed ones, on the same tree
                    for (int r = 0; r <= reps && !rc; r++) {
                        if (r == 1) (void)hipEventRecord(s->gev[0], s->stream);
                        hipLaunchKernelGGL(k_walk_reset, dim3(1), dim3(64), 0, s->stream, meta);
                        rc = enqueue_walk(s, s->walk, a, false);
                    }
                    if (rc) break;
                    (void)hipEventRecord(s->gev[1], s->stream);
                    HIPCHK(hipMemcpyAsync(s->hmeta, meta, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
                    HIPCHK(hipStreamSynchronize(s->stream));
                    const Meta& h = *s->hmeta;
                    if (h.flags & kFlagListOverflow) {  // this group size needs longer lists: grow, time again
                        (void)regrow(s);
                        continue;
                    }
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, s->gev[0], s->gev[1]));
                    gx_tune_entry e;
                    e.group_size = gsize;
                    e.persistent = pers;
                    e.stack = kWalkConfig[cfg].stack;
                    e.parts = kWalkConfig[cfg].parts;
                    e.waves = kWalkConfig[cfg].waves;
                    e.ms = (h.flags & kFlagStackOverflow) ? INFINITY : ms / (float)reps;  // a stack overflow disqualifies
                    e.pp_interactions = (int64_t)h.pp - (int64_t)h.ntargets;
                    e.pc_interactions = (int64_t)h.pc;
                    t.push_back(e);
                    done = true;
                }
            }
        if (rc) break;
    }
    s->timing = timing;
    s->lastEvents = kNoEvents;
    if (rc) return rc;
    size_t best = 0;
    for (size_t i = 1; i < t.size(); i++)
        if (t[i].ms < t[best].ms) best = i;
    if (t.empty() || !(t[best].ms < INFINITY)) return set_err(GX_ERR_CAPACITY, "no walk candidate completed");
    s->gsize = t[best].group_size;
    s->persistent = t[best].persistent;
    for (int c = 0; c < kWalkConfigs; c++)
        if (kWalkConfig[c].stack == t[best].stack && kWalkConfig[c].parts == t[best].parts &&
            kWalkConfig[c].waves == t[best].waves)
            s->walkConfig = c;
    s->tuned.push_back({a.n, count, s->walk, s->monopole, s->gsize, s->persistent, s->walkConfig, t});
    report(t);
    // leave the forces of the chosen walk in ax, ay, az
    return run_dev(s, a);
}

int gx_graph_capture_step(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                          const double* mass, const uint8_t* kind, const uint8_t* alive, const int64_t* offsets,
                          int32_t count, const double* theta, const double* eps, int dark_matter_off, double dt,
                          double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    int rc;
    if ((rc = check_batch(offsets, count, theta, eps))) return rc;
    if (s->stream == nullptr) return set_err(GX_ERR_ARG, "capture needs a non-null stream");
    const long long n = offsets[count];
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    if ((rc = check_arrays(a)) || (rc = set_instances(s, count, offsets, theta, 0, eps, 0))) return rc;
    // an eager force evaluation sizes every arena for the captured step
    if ((rc = run_dev(s, a))) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->graphExec) HIPCHK(hipGraphExecDestroy(s->graphExec));
    if (s->graph) HIPCHK(hipGraphDestroy(s->graph));
    s->graphExec = nullptr;
    s->graph = nullptr;
    const int timing = s->timing;
    s->timing = 0;  // no event records inside the graph
    s->capturing = true;
    s->allLevels = true;  // replays cannot repeat a step with more levels
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    rc = kick(s, vx, vy, vz, ax, ay, az, n, 0.5 * dt);
    if (!rc) rc = drift(s, x, y, z, vx, vy, vz, n, dt);
    if (!rc) rc = enqueue_build(s, a);
    if (!rc) rc = enqueue_walk(s, s->walk, a);
    if (!rc) rc = kick(s, vx, vy, vz, ax, ay, az, n, 0.5 * dt);
    hipError_t mc = hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(s->stream, &g);
    s->timing = timing;
    s->capturing = false;
    s->allLevels = false;
    s->pending = false;
    s->lastEvents = kNoEvents;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (mc != hipSuccess || ec != hipSuccess)
        return set_err(GX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(mc != hipSuccess ? mc : ec));
    s->graph = g;
    HIPCHK(hipGraphInstantiate(&s->graphExec, g, nullptr, nullptr, 0));
    s->graphBufs = s->buffer_ptrs();
    s->lastWalk = s->walk;
    return GX_OK;
}

int gx_graph_launch(gx_solver* s, int32_t steps) {
    if (!s || !s->graphExec || steps < 0) return set_err(GX_ERR_ARG, "no captured step");
    // a later eager call that needed larger buffers freed the ones the graph uses
    if (s->buffer_ptrs() != s->graphBufs)
        return set_err(GX_ERR_ARG, "the captured step's buffers were reallocated by a larger step; capture again");
    if (steps == 0) return GX_OK;
    // an unsynchronised earlier run: its flags stay in the sticky word and are
    // reported with this run's
    HIPCHK(hipEventRecord(s->gev[0], s->stream));
    for (int i = 0; i < steps; i++) HIPCHK(hipGraphLaunch(s->graphExec, s->stream));
    HIPCHK(hipEventRecord(s->gev[1], s->stream));
    s->graphSteps = steps;
    s->lastEvents = kGraphEvents;
    s->pending = true;
    return GX_OK;
}

int gx_potential_energy_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                            const uint8_t* kind, const uint8_t* alive, int64_t n, double eps, int dark_matter_off,
                            double c, double* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    if (n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (s->pending) (void)finish_step(s);
    *out = 0;
    if (n == 0) return GX_OK;
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, nullptr, nullptr, nullptr);
    int rc;
    if ((rc = set_single(s, n, std::min(theta_of(s->theta), 0.9), eps))) return rc;  // diagTheta, bh.go:20-23
    for (int attempt = 0; attempt < 6; attempt++) {
        if ((rc = enqueue_build(s, a))) return rc;
        // n star partials, then 2 per hole (the holes are at most n)
        if ((rc = ensure(s->partial, (size_t)n * 8 * 3))) return rc;
        double* part = (double*)s->partial.p;
        hipLaunchKernelGGL(k_potential64, dim3(blocks(n, 256)), dim3(256), 0, s->stream, (const double4*)s->spos.p,
                           (const TravNode*)s->trav.p, (const Mult64*)s->m64.p, (Meta*)s->meta.p, make_batch(s),
                           !s->monopole, (const int32_t*)s->order.p, (const int32_t*)s->holes.p, x, y, z, mass, kind,
                           dark_matter_off, c, part);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (s->hmeta->flags & kFlagNodeOverflow) {
            s->capNodes *= 2;
            continue;
        }
        if (s->hmeta->flags & kFlagDepth) {
            s->levels = kMortonBits;
            continue;
        }
        if (s->hmeta->flags & kFlagTopology)
            return set_err(GX_ERR_PIPELINE, "tree links failed their check (0x%x at node %lld of %lld)",
                           s->hmeta->topoWhy, s->hmeta->topoNode, s->hmeta->nnodes);
        const long long ns = s->hmeta->nstars, nh = s->hmeta->nholes;
        if (nh > 0) {
            hipLaunchKernelGGL(k_hole_potential, dim3((unsigned)nh), dim3(256), 0, s->stream, (const double4*)s->spos.p,
                               (const Meta*)s->meta.p, (const int32_t*)s->holes.p, x, y, z, mass, eps * eps, c,
                               part + ns);
            HIPCHK(hipGetLastError());
        }
        // fixed-order sums on the host, as potential.go:20-39: stars in sorted
        // order, then per hole its star pairs and its hole pairs
        std::vector<double> h((size_t)(ns + 2 * nh));
        if (!h.empty()) HIPCHK(hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        double sum = 0;
        for (long long k = 0; k < ns; k++) sum += h[k];
        for (long long g = 0; g < nh; g++) {
            sum -= h[ns + 2 * g];
            sum -= h[ns + 2 * g + 1];
        }
        *out = -sum / 2;
        return GX_OK;
    }
    return set_err(GX_ERR_CAPACITY, "node arena did not fit after regrowing");
}

int gx_direct_sample_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                         const uint8_t* kind, int64_t n, const int64_t* targets, int64_t nt, double eps,
                         int dark_matter_off, double* out) {
    if (!s || !out || (nt > 0 && !targets)) return set_err(GX_ERR_ARG, "null argument");
    if (n < 0 || nt < 0 || nt > (1LL << 30)) return set_err(GX_ERR_ARG, "bad sizes");
    if (nt == 0) return GX_OK;
    if (s->pending) (void)finish_step(s);
    const int tblocks = (int)((nt + 255) / 256);
    int splits = std::max(1, 2048 / tblocks);
    splits = (int)std::min<long long>(splits, std::max(1LL, (long long)(n / 1024)));
    int rc;
    if ((rc = ensure(s->partial, (size_t)splits * nt * 3 * 8))) return rc;
    hipLaunchKernelGGL(k_direct_sample, dim3(tblocks, splits), dim3(256), 0, s->stream, x, y, z, mass, kind,
                       (long long)n, (const long long*)targets, (int)nt, eps * eps, dark_matter_off, splits,
                       (double*)s->partial.p);
    hipLaunchKernelGGL(k_direct_reduce, dim3(tblocks), dim3(256), 0, s->stream, (const double*)s->partial.p, (int)nt,
                       splits, out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

int gx_potential_exact_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                           const uint8_t* kind, int64_t n, double eps, int dark_matter_off, double c, double* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    if (n < 0 || n > (1LL << 31) - 1) return set_err(GX_ERR_ARG, "bad n");
    if (n > 0 && (!x || !y || !z || !mass || !kind)) return set_err(GX_ERR_ARG, "null array");
    if (s->pending) (void)finish_step(s);
    *out = 0;
    if (n < 2) return GX_OK;
    const long long rowBlocks = (n + 255) / 256;
    // enough workgroups to fill the device; slices of at least 1,024 columns
    int slices = (int)std::max(1LL, std::min<long long>(4096 / rowBlocks + 1, (n + 1023) / 1024));
    int rc;
    if ((rc = ensure(s->partial, (size_t)slices * n * 8))) return rc;
    hipLaunchKernelGGL(k_potential_exact, dim3((unsigned)rowBlocks, (unsigned)slices), dim3(256), 0, s->stream, x, y,
                       z, mass, kind, (long long)n, eps * eps, dark_matter_off, c, slices, (double*)s->partial.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h((size_t)slices * n);
    HIPCHK(hipMemcpyAsync(h.data(), s->partial.p, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    // fixed order: rows in state order, each row's slices in column order
    double sum = 0;
    for (long long i = 0; i < n; i++) {
        double row = 0;
        for (int k = 0; k < slices; k++) row += h[(size_t)k * n + i];
        sum += row;
    }
    *out = sum;
    return GX_OK;
}

double gx_last_seconds(const gx_solver* s) { return s ? s->lastSeconds : 0; }

int gx_get_stats(const gx_solver* s, gx_stats* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    *out = s->stats;
    out->walk = s->walk;
    return GX_OK;
}

int64_t gx_export(gx_solver* s, int what, void* host, int64_t count) {
    if (!s || !host || count < 0) return set_err(GX_ERR_ARG, "bad argument"), -1;
    if (s->pending) (void)finish_step(s);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return -1;
    const Meta& h = *s->hmeta;
    const void* src = nullptr;
    long long avail = 0;
    size_t item = 0;
    switch (what) {
        case 0: src = s->keys.p; avail = h.nstars; item = 8; break;
        case 1: src = s->order.p; avail = h.nstars; item = 4; break;
        case 2: src = s->spos.p; avail = h.nstars; item = 32; break;
        case 3: src = s->trav.p; avail = std::min(h.nnodes, s->capNodes); item = sizeof(TravNode); break;
        case 4: src = s->m64.p; avail = std::min(h.nnodes, s->capNodes); item = sizeof(Mult64); break;
        default: return set_err(GX_ERR_ARG, "unknown export %d", what), -1;
    }
    long long k = std::min<long long>(avail, count);
    if (k > 0 && hipMemcpy(host, src, (size_t)k * item, hipMemcpyDeviceToHost) != hipSuccess)
        return set_err(GX_ERR_HIP, "export copy failed"), -1;
    return k;
}

int gx_phase_ms(const gx_solver* s, float* ms, int n) {
    if (!s || !ms) return 0;
    int k = std::min(n, kPhases);
    for (int i = 0; i < k; i++) ms[i] = s->phase[i];
    return k;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// Force step without allocation: every buffer of the step is a slice of one
// static device array per GPU (gx_accelerations_noalloc).
// ---------------------------------------------------------------------------
namespace {

constexpr long long kNaBodies = GX_NOALLOC_MAX_BODIES;
constexpr long long kNaNodes = kNaBodies + kNaBodies / 4;  // clustered trees need more than 0.2 nodes per star
constexpr int kNaSlots = 256;                             // list regions of the work queue
constexpr int kNaFull = 3072, kNaPart = 3072, kNaLeaf = 768;  // fixed list capacities per region
constexpr long long kNaTiles = (kNaBodies + kSortTile - 1) / kSortTile;
constexpr long long kNaCollect = (kNaBodies + kCollectTile - 1) / kCollectTile;

constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }

// slice sizes, in the order they are carved
constexpr size_t kNaSizes[] = {
    na_round(8 * kNaBodies), na_round(8 * kNaBodies),    // keys, keys2
    na_round(4 * kNaBodies), na_round(4 * kNaBodies),    // order, order2
    na_round(4 * kNaBodies), na_round(4 * kNaBodies),    // rank, holes
    na_round(256 * 4 * kNaTiles), na_round(256 * 4),     // hist, digitTotal
    na_round(4 * kNaCollect), na_round(4 * kNaCollect),  // blockStar, blockHole
    na_round(48 * kNaCollect),                           // blockBox
    na_round(32 * kNaBodies), na_round(sizeof(Meta)),    // spos, meta
    na_round(sizeof(TravNode) * kNaNodes), na_round(sizeof(Mult64) * kNaNodes),  // trav, m64
    na_round(sizeof(Mult32) * kNaNodes), na_round(sizeof(Mult16) * kNaNodes),    // m32, m16
    na_round(4ull * kNaSlots * kNaFull), na_round(12ull * kNaSlots * kNaPart),    // lfull, lpart
    na_round(12ull * kNaSlots * kNaLeaf), na_round(sizeof(GroupCounts) * kNaSlots),  // lleaf, lcounts
    256, 256, 256, 256, 256, 256, 256,  // instCfg, boxEnc, starCount, starOff, groupOff, root, half (one instance)
};
constexpr size_t na_total() {
    size_t t = 0;
    for (size_t b : kNaSizes) t += b;
    return t;
}
static_assert(na_total() <= (size_t)GX_NOALLOC_SCRATCH_BYTES, "the no-allocation scratch is too small");

}  // namespace

// The scratch: one copy per GPU, held for the life of the process.  Host code
// never uses this name as an address; it asks each device for the array's
// address (hipGetSymbolAddress) and uses that.
static __device__ __align__(256) unsigned char scratch[GX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNaDevices = 16;

struct NoAllocDevice {
    std::once_flag once;
    int rc = GX_OK;
    char msg[256] = "";
    std::mutex lock;  // calls on one device are serialised on the host
    gx_solver s;
    Meta hmeta;
    alignas(64) char inst[256];
};
NoAllocDevice g_na[kNaDevices];

// Carves the device's scratch into the solver's buffers (once per device,
// with that device current).
void na_setup(NoAllocDevice& d) {
    unsigned char* base = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&base, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !base) {
        d.rc = GX_ERR_HIP;
        snprintf(d.msg, sizeof d.msg, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        return;
    }
    gx_solver& s = d.s;
    DevBuf* order[] = {&s.keys, &s.keys2, &s.order, &s.order2, &s.rank, &s.holes, &s.hist, &s.digitTotal,
                       &s.blockStar, &s.blockHole, &s.blockBox, &s.spos, &s.meta, &s.trav, &s.m64, &s.m32, &s.m16,
                       &s.lfull, &s.lpart, &s.lleaf, &s.lcounts, &s.instCfg, &s.boxEnc, &s.starCount,
                       &s.starOff, &s.groupOff, &s.root, &s.half};
    static_assert(sizeof(order) / sizeof(order[0]) == sizeof(kNaSizes) / sizeof(kNaSizes[0]), "one size per slice");
    size_t off = 0;
    for (size_t i = 0; i < sizeof(kNaSizes) / sizeof(kNaSizes[0]); i++) {
        order[i]->p = base + off;
        order[i]->bytes = kNaSizes[i];
        off += kNaSizes[i];
    }
    s.hmeta = &d.hmeta;
    s.instHost = d.inst;
    s.instHostBytes = sizeof d.inst;
    s.instCopied = nullptr;  // no upload cache
    s.capturing = true;      // no events
    s.fixedCaps = true;
    s.allLevels = true;      // no repeat for a deeper tree
    s.persistent = 2;
    s.maxSlots = kNaSlots;
    s.cf = kNaFull;
    s.cp = kNaPart;
    s.cl = kNaLeaf;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&s.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        s.cus = 1;
    for (int c = 0; c < kWalkConfigs; c++) s.blocksPerCu[c] = std::max(walk32_blocks_per_cu(c), 1);
    // the step metadata starts at zero (its sticky flags above all)
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, nullptr, (unsigned long long*)s.meta.p,
                       (int)(sizeof(Meta) / 8), 0ull);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        d.rc = GX_ERR_HIP;
        snprintf(d.msg, sizeof d.msg, "scratch set-up: %s", hipGetErrorString(e));
    }
}

struct NoAllocScope {
    NoAllocScope() { t_noAlloc = true; }
    ~NoAllocScope() { t_noAlloc = false; }
};

}  // namespace

extern "C" {

int gx_accelerations_noalloc(const double* x, const double* y, const double* z, const double* mass,
                             const uint8_t* kind, const uint8_t* alive, int64_t n, double theta, double eps,
                             int dark_matter_off, int walk, double* ax, double* ay, double* az, void* stream) {
    if (n < 0 || n > kNaBodies)
        return set_err(GX_ERR_ARG, "n = %lld: the no-allocation entry point takes at most %lld bodies", (long long)n,
                       kNaBodies);
    if (!(theta >= 0) || walk < GX_WALK_FP32 || walk > GX_WALK_FP32_Q16)
        return set_err(GX_ERR_ARG, "bad theta %g or walk %d", theta, walk);
    int rc = device_ready();
    if (rc) return rc;
    if (n == 0) return GX_OK;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaDevices) return set_err(GX_ERR_ARG, "device %d beyond the %d with a scratch", dev, kNaDevices);
    NoAllocDevice& d = g_na[dev];
    std::call_once(d.once, [&d] { na_setup(d); });
    if (d.rc) return set_err(d.rc, "%s", d.msg);
    std::lock_guard<std::mutex> guard(d.lock);
    NoAllocScope scope;
    gx_solver* s = &d.s;
    s->stream = (hipStream_t)stream;
    s->theta = theta;
    s->monopole = 0;
    s->walk = walk;
    s->timing = 0;
    s->pending = false;
    const StepArgs a = step_args(x, y, z, mass, kind, alive, n, dark_matter_off, ax, ay, az);
    if ((rc = check_arrays(a)) || (rc = set_single(s, n, theta, eps))) return rc;
    return run_dev(s, a);
}

}  // extern "C"
