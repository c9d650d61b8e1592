// gx_api.hip -- C ABI and step orchestration (host side of the backend).
//
// One call of the force step enqueues, on one stream and with no host
// synchronisation in between: collect (3 launches), keys, 8 radix passes
// (3 launches each), gather, one topology launch per tree level, one moment
// launch per level, the walk and the hole fix-up.  Sizes that only the device
// knows (number of tree stars, nodes per level) are read by the kernels from
// the Meta record; grids are sized from the body count n.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

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

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
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

int ensure(DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return GX_OK;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    HIPCHK(hipMalloc(&b.p, want));
    b.bytes = want;
    return GX_OK;
}

constexpr int kPhases = 9;
constexpr int kTreeGrid = 1024;

}  // namespace

struct gx_solver {
    double theta = 0;
    int monopole = 0;
    int walk = GX_WALK_FP32;
    int timing = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    // host-API staging
    DevBuf x, y, z, m, kind, alive, ax, ay, az;
    // arena
    DevBuf keys, keys2, order, order2, holes, hist, digitTotal, blockStar, blockHole, blockBox, spos;
    DevBuf trav, m64, m32, m16, meta, partial, targets;
    DevBuf lfull, lpart, lleaf, lcounts;
    int cf = 2048, cp = 512, cl = 256;  // FP32-walk list capacities per group (grow on demand)
    long long capNodes = 0;
    Meta* hmeta = nullptr;  // pinned mirror of the device Meta
    hipEvent_t ev[kPhases + 1] = {};
    float phase[kPhases] = {};
    double lastSeconds = 0;
    gx_stats stats = {};
    long long fallbacks = 0;
    // the pending step (for gx_sync)
    bool pending = false;
    long long pendingN = 0;
    int lastWalk = GX_WALK_FP32;
    // FP32 walk as two kernels (traversal, evaluation) instead of one; set
    // from GX_SPLIT_WALK=1 for profiling the two phases separately.
    bool splitWalk = false;

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

gx_solver* gx_solver_new(void) {
    if (gx_init(nullptr, 0) != GX_OK) {
        set_err(g_init_code, "%s", g_init_msg);
        return nullptr;
    }
    gx_solver* s = new gx_solver();
    if (hipStreamCreateWithFlags(&s->own, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&s->hmeta, sizeof(Meta), hipHostMallocDefault) != hipSuccess) {
        set_err(GX_ERR_HIP, "stream or pinned allocation failed");
        delete s;
        return nullptr;
    }
    memset(s->hmeta, 0, sizeof(Meta));
    s->stream = s->own;
    for (auto& e : s->ev) (void)hipEventCreate(&e);
    const char* sw = getenv("GX_SPLIT_WALK");
    s->splitWalk = sw && sw[0] == '1';

    return s;
}

void gx_solver_free(gx_solver* s) {
    if (!s) return;
    hipStreamSynchronize(s->stream);
    DevBuf* bufs[] = {&s->x,     &s->y,         &s->z,         &s->m,        &s->kind,      &s->alive,
                      &s->ax,    &s->ay,        &s->az,        &s->keys,     &s->keys2,     &s->order,
                      &s->order2, &s->holes,    &s->hist,      &s->digitTotal, &s->blockStar, &s->blockHole,
                      &s->blockBox, &s->spos,   &s->trav,      &s->m64,      &s->m32,  &s->m16,     &s->meta,
                      &s->partial, &s->targets, &s->lfull, &s->lpart, &s->lleaf, &s->lcounts};
    for (DevBuf* b : bufs)
        if (b->p) hipFree(b->p);
    for (auto& e : s->ev)
        if (e) hipEventDestroy(e);
    if (s->hmeta) hipHostFree(s->hmeta);
    if (s->own) hipStreamDestroy(s->own);
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

int gx_solver_set_stream(gx_solver* s, void* stream) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    s->stream = stream ? (hipStream_t)stream : s->own;
    return GX_OK;
}

}  // extern "C"

namespace {

double theta_of(const gx_solver* s) { return s->theta == 0 ? 0.9 : s->theta; }

int reserve(gx_solver* s, long long n) {
    const long long nb = (n + kCollectTile - 1) / kCollectTile;
    const long long nt = (n + kSortTile - 1) / kSortTile;
    int rc;
    if ((rc = ensure(s->keys, n * 8)) || (rc = ensure(s->keys2, n * 8)) || (rc = ensure(s->order, n * 4)) ||
        (rc = ensure(s->order2, n * 4)) || (rc = ensure(s->holes, n * 4)) ||
        (rc = ensure(s->hist, (size_t)256 * std::max(nt, 1LL) * 4)) || (rc = ensure(s->digitTotal, 256 * 4)) ||
        (rc = ensure(s->blockStar, std::max(nb, 1LL) * 4)) || (rc = ensure(s->blockHole, std::max(nb, 1LL) * 4)) ||
        (rc = ensure(s->blockBox, std::max(nb, 1LL) * 48)) || (rc = ensure(s->spos, std::max(n, 1LL) * 32)) ||
        (rc = ensure(s->meta, sizeof(Meta))))
        return rc;
    long long wantNodes = std::max(4096LL, n);
    if (s->capNodes < wantNodes) s->capNodes = wantNodes;
    if ((rc = ensure(s->trav, (size_t)s->capNodes * sizeof(TravNode))) ||
        (rc = ensure(s->m64, (size_t)s->capNodes * sizeof(Mult64))) ||
        (rc = ensure(s->m32, (size_t)s->capNodes * sizeof(Mult32))) ||
        (rc = ensure(s->m16, (size_t)s->capNodes * sizeof(Mult16))))
        return rc;
    // ensure() may have over-allocated: use all of it
    s->capNodes = (long long)std::min({s->trav.bytes / sizeof(TravNode), s->m64.bytes / sizeof(Mult64),
                                       s->m32.bytes / sizeof(Mult32), s->m16.bytes / sizeof(Mult16)});
    return GX_OK;
}

inline unsigned blocks(long long n, int per) { return (unsigned)std::max(1LL, (n + per - 1) / per); }

void mark(gx_solver* s, int i) {
    if (s->timing) hipEventRecord(s->ev[i], s->stream);
}

// Build the tree for the device arrays (collect .. moments).  theta is baked
// into crit2 (build.go:209-211).
int enqueue_build(gx_solver* s, const double* x, const double* y, const double* z, const double* m,
                  const uint8_t* kind, const uint8_t* alive, long long n, int dmOff, double theta) {
    int rc = reserve(s, n);
    if (rc) return rc;
    hipStream_t st = s->stream;
    Meta* meta = (Meta*)s->meta.p;
    const long long nb = (n + kCollectTile - 1) / kCollectTile;
    const int ntiles = (int)((n + kSortTile - 1) / kSortTile);
    hipEventRecord(s->ev[0], st);
    hipLaunchKernelGGL(k_collect_count, dim3(blocks(n, kCollectTile)), dim3(256), 0, st, x, y, z, kind, alive, n,
                       (int*)s->blockStar.p, (int*)s->blockHole.p, (double*)s->blockBox.p);
    hipLaunchKernelGGL(k_collect_scan, dim3(1), dim3(1024), 0, st, (int)std::max(nb, 1LL), (int*)s->blockStar.p,
                       (int*)s->blockHole.p, (const double*)s->blockBox.p, meta);
    hipLaunchKernelGGL(k_collect_scatter, dim3(blocks(n, kCollectTile)), dim3(256), 0, st, kind, alive, n,
                       (const int*)s->blockStar.p, (const int*)s->blockHole.p, (int32_t*)s->order.p,
                       (int32_t*)s->holes.p);
    mark(s, 1);
    hipLaunchKernelGGL(k_keys, dim3(blocks(n, 256)), dim3(256), 0, st, x, y, z, (const int32_t*)s->order.p, meta,
                       (uint64_t*)s->keys.p);
    mark(s, 2);
    uint64_t* kin = (uint64_t*)s->keys.p;
    uint64_t* kout = (uint64_t*)s->keys2.p;
    int32_t* vin = (int32_t*)s->order.p;
    int32_t* vout = (int32_t*)s->order2.p;
    for (int pass = 0; pass < 8; pass++) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_sort_hist, dim3(std::max(ntiles, 1)), dim3(256), 0, st, kin, meta, shift,
                           (uint32_t*)s->hist.p, std::max(ntiles, 1));
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(256), 0, st, (uint32_t*)s->hist.p, std::max(ntiles, 1),
                           (uint32_t*)s->digitTotal.p);
        hipLaunchKernelGGL(k_sort_scatter, dim3(std::max(ntiles, 1)), dim3(256), 0, st, kin, vin, kout, vout, meta,
                           shift, (const uint32_t*)s->hist.p, (const uint32_t*)s->digitTotal.p, std::max(ntiles, 1));
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    mark(s, 3);
    hipLaunchKernelGGL(k_gather, dim3(blocks(n, 256)), dim3(256), 0, st, x, y, z, m, kind,
                       (const int32_t*)s->order.p, meta, dmOff, (double4*)s->spos.p);
    mark(s, 4);
    hipLaunchKernelGGL(k_tree_init, dim3(1), dim3(64), 0, st, meta, (TravNode*)s->trav.p);
    for (int L = 0; L <= kMortonBits; L++)
        hipLaunchKernelGGL(k_topo_level, dim3(kTreeGrid), dim3(256), 0, st, L, (const uint64_t*)s->keys.p,
                           (TravNode*)s->trav.p, meta, s->capNodes);
    mark(s, 5);
    for (int L = kMortonBits; L >= 0; L--)
        hipLaunchKernelGGL(k_moments_level, dim3(kTreeGrid), dim3(256), 0, st, L, (const double4*)s->spos.p,
                           (TravNode*)s->trav.p, (Mult64*)s->m64.p, (Mult32*)s->m32.p, (Mult16*)s->m16.p, meta,
                           theta);
    hipLaunchKernelGGL(k_tree_finish, dim3(1), dim3(64), 0, st, meta);
    mark(s, 6);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

int enqueue_walk(gx_solver* s, int walk, long long n, double eps, double* ax, double* ay, double* az) {
    hipStream_t st = s->stream;
    Meta* meta = (Meta*)s->meta.p;
    const double eps2 = eps * eps;
    if (walk == GX_WALK_FP32 || walk == GX_WALK_FP32_Q16) {
        const long long ng = std::max(1LL, (n + 63) / 64);
        int rc;
        if ((rc = ensure(s->lfull, (size_t)ng * s->cf * 4)) || (rc = ensure(s->lpart, (size_t)ng * s->cp * 16)) ||
            (rc = ensure(s->lleaf, (size_t)ng * s->cl * 16)) || (rc = ensure(s->lcounts, (size_t)ng * 16)))
            return rc;
        Lists lists;
        lists.full = (uint32_t*)s->lfull.p;
        lists.part = (uint4*)s->lpart.p;
        lists.leaf = (uint4*)s->lleaf.p;
        lists.counts = (int4*)s->lcounts.p;
        lists.cf = s->cf;
        lists.cp = s->cp;
        lists.cl = s->cl;
        launch_walk32((const double4*)s->spos.p, (const TravNode*)s->trav.p, (const Mult32*)s->m32.p,
                      walk == GX_WALK_FP32_Q16 ? (const Mult16*)s->m16.p : nullptr, (const int32_t*)s->order.p, meta,
                      (n + 63) / 64, (float)eps2, !s->monopole, lists, ax, ay, az, st, s->timing ? s->ev[7] : nullptr,
                      s->splitWalk);
    } else {
        hipLaunchKernelGGL(k_walk64, dim3(blocks(n, 256)), dim3(256), 0, st, (const double4*)s->spos.p,
                           (const TravNode*)s->trav.p, (const Mult64*)s->m64.p, (const int32_t*)s->order.p, meta,
                           eps2, !s->monopole, ax, ay, az);
    }
    hipLaunchKernelGGL(k_zero_holes, dim3(64), dim3(256), 0, st, (const int32_t*)s->holes.p, meta, ax, ay, az);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

int enqueue_step(gx_solver* s, const double* x, const double* y, const double* z, const double* m,
                 const uint8_t* kind, const uint8_t* alive, long long n, double eps, int dmOff, double* ax,
                 double* ay, double* az) {
    if (n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (n >= (1LL << 31) - 1) return set_err(GX_ERR_ARG, "n = %lld exceeds the int32 index range", n);
    int rc = enqueue_build(s, x, y, z, m, kind, alive, n, dmOff, theta_of(s));
    if (rc) return rc;
    if ((rc = enqueue_walk(s, s->walk, n, eps, ax, ay, az))) return rc;
    hipEventRecord(s->ev[kPhases], s->stream);
    s->lastWalk = s->walk;
    HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
    s->pending = true;
    s->pendingN = n;
    return GX_OK;
}

// Waits for the pending step; fills stats and phase times.
int finish_step(gx_solver* s) {
    if (!s->pending) return GX_OK;
    s->pending = false;
    HIPCHK(hipStreamSynchronize(s->stream));
    const Meta& h = *s->hmeta;
    float total = 0;
    hipEventElapsedTime(&total, s->ev[0], s->ev[kPhases]);
    s->lastSeconds = total * 1e-3;
    if (s->timing) {
        for (int i = 0; i < 6; i++) hipEventElapsedTime(&s->phase[i], s->ev[i], s->ev[i + 1]);
        hipEventElapsedTime(&s->phase[6], s->ev[6], s->ev[kPhases]);
        s->phase[7] = total;
        s->phase[8] = 0;
        if (s->lastWalk != GX_WALK_FP64 && s->splitWalk) {  // traversal | evaluation split
            float trav = 0;
            hipEventElapsedTime(&trav, s->ev[6], s->ev[7]);
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
    s->stats.pp_interactions = (int64_t)h.pp - (int64_t)h.nstars;
    s->stats.max_stack = h.maxStack;
    s->stats.flags = h.flags;
    s->stats.depth = depth;
    s->stats.fallbacks = s->fallbacks;
    s->stats.list_entries = (int64_t)h.listEntries;
    s->stats.rounds = (int64_t)h.rounds;
    s->stats.mixed_nodes = (int64_t)h.mixed;
    for (int i = 0; i < 3; i++) s->stats.max_list[i] = h.maxList[i];
    s->stats.max_list[3] = 0;
    if (h.flags & kFlagNodeOverflow)
        return set_err(GX_ERR_CAPACITY, "node arena of %lld nodes too small", s->capNodes);
    return GX_OK;
}

int run_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* m, const uint8_t* kind,
            const uint8_t* alive, long long n, double eps, int dmOff, double* ax, double* ay, double* az) {
    for (int attempt = 0; attempt < 6; attempt++) {
        int rc = enqueue_step(s, x, y, z, m, kind, alive, n, eps, dmOff, ax, ay, az);
        if (rc) return rc;
        rc = finish_step(s);
        if (rc == GX_ERR_CAPACITY) {
            s->capNodes *= 2;
            continue;
        }
        if (rc) return rc;
        if (s->hmeta->flags & kFlagListOverflow) {
            // some group's interaction list did not fit: grow and redo the walk
            const Meta& h = *s->hmeta;
            s->cf = std::max(s->cf, (int)(h.maxList[0] * 1.25) + 64);
            s->cp = std::max(s->cp, (int)(h.maxList[1] * 1.25) + 64);
            s->cl = std::max(s->cl, (int)(h.maxList[2] * 1.25) + 64);
            continue;
        }
        if (s->hmeta->flags & kFlagStackOverflow) {
            // the FP32 walk's LDS stack was too small for this tree: redo the
            // walk with the FP64 per-particle walk (same interaction set)
            s->fallbacks++;
            if ((rc = enqueue_walk(s, GX_WALK_FP64, n, eps, ax, ay, az))) return rc;
            HIPCHK(hipStreamSynchronize(s->stream));
            s->stats.fallbacks = s->fallbacks;
        }
        return GX_OK;
    }
    return set_err(GX_ERR_CAPACITY, "node arena did not fit after regrowing");
}

}  // namespace

extern "C" {

int gx_accelerations_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                         const uint8_t* kind, const uint8_t* alive, int64_t n, double eps, int dark_matter_off,
                         double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (n > 0 && (!x || !y || !z || !mass || !kind || !alive || !ax || !ay || !az))
        return set_err(GX_ERR_ARG, "null array");
    if (s->pending) finish_step(s);
    return run_dev(s, x, y, z, mass, kind, alive, n, eps, dark_matter_off, ax, ay, az);
}

int gx_enqueue_accelerations_dev(gx_solver* s, const double* x, const double* y, const double* z,
                                 const double* mass, const uint8_t* kind, const uint8_t* alive, int64_t n,
                                 double eps, int dark_matter_off, double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (n > 0 && (!x || !y || !z || !mass || !kind || !alive || !ax || !ay || !az))
        return set_err(GX_ERR_ARG, "null array");
    s->pending = false;  // a new step supersedes an unsynchronised one
    return enqueue_step(s, x, y, z, mass, kind, alive, n, eps, dark_matter_off, ax, ay, az);
}

int gx_sync(gx_solver* s) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    if (!s->pending) {
        HIPCHK(hipStreamSynchronize(s->stream));
        return GX_OK;
    }
    int rc = finish_step(s);
    if (rc) return rc;
    if (s->hmeta->flags & (kFlagStackOverflow | kFlagListOverflow))
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

int gx_kick_dev(gx_solver* s, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                const double* az, int64_t n, double h) {
    if (!s || n < 0) return set_err(GX_ERR_ARG, "bad argument");
    if (n == 0) return GX_OK;
    hipLaunchKernelGGL(k_kick, dim3((unsigned)std::min<long long>(blocks(n, 256), 4096)), dim3(256), 0, s->stream,
                       (long long)n, vx, vy, vz, ax, ay, az, h);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

int gx_drift_dev(gx_solver* s, double* x, double* y, double* z, const double* vx, const double* vy,
                 const double* vz, int64_t n, double dt) {
    if (!s || n < 0) return set_err(GX_ERR_ARG, "bad argument");
    if (n == 0) return GX_OK;
    hipLaunchKernelGGL(k_drift, dim3((unsigned)std::min<long long>(blocks(n, 256), 4096)), dim3(256), 0, s->stream,
                       (long long)n, x, y, z, vx, vy, vz, dt);
    HIPCHK(hipGetLastError());
    return GX_OK;
}

int gx_step_dev(gx_solver* s, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                const double* mass, const uint8_t* kind, const uint8_t* alive, int64_t n, double eps,
                int dark_matter_off, double dt, int acc_valid, double* ax, double* ay, double* az) {
    if (!s) return set_err(GX_ERR_ARG, "null solver");
    int rc;
    if (!acc_valid &&
        (rc = gx_accelerations_dev(s, x, y, z, mass, kind, alive, n, eps, dark_matter_off, ax, ay, az)))
        return rc;
    if ((rc = gx_kick_dev(s, vx, vy, vz, ax, ay, az, n, 0.5 * dt))) return rc;
    if ((rc = gx_drift_dev(s, x, y, z, vx, vy, vz, n, dt))) return rc;
    if ((rc = gx_accelerations_dev(s, x, y, z, mass, kind, alive, n, eps, dark_matter_off, ax, ay, az))) return rc;
    if ((rc = gx_kick_dev(s, vx, vy, vz, ax, ay, az, n, 0.5 * dt))) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return GX_OK;
}

int gx_potential_energy_dev(gx_solver* s, const double* x, const double* y, const double* z, const double* mass,
                            const uint8_t* kind, const uint8_t* alive, int64_t n, double eps, int dark_matter_off,
                            double* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    if (n < 0) return set_err(GX_ERR_ARG, "negative n");
    if (s->pending) finish_step(s);
    *out = 0;
    if (n == 0) return GX_OK;
    const double theta = std::min(theta_of(s), 0.9);  // diagTheta, bh.go:20-23
    for (int attempt = 0; attempt < 6; attempt++) {
        int rc = enqueue_build(s, x, y, z, mass, kind, alive, n, dark_matter_off, theta);
        if (rc) return rc;
        if ((rc = ensure(s->partial, (size_t)n * 8))) return rc;
        hipLaunchKernelGGL(k_potential64, dim3(blocks(n, 256)), dim3(256), 0, s->stream, (const double4*)s->spos.p,
                           (const TravNode*)s->trav.p, (const Mult64*)s->m64.p, (Meta*)s->meta.p, eps * eps,
                           !s->monopole, (double*)s->partial.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(s->hmeta, s->meta.p, sizeof(Meta), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (s->hmeta->flags & kFlagNodeOverflow) {
            s->capNodes *= 2;
            continue;
        }
        if (s->hmeta->nholes > 0)
            return set_err(GX_ERR_UNSUPPORTED, "potential energy with black holes is not implemented");
        const long long ns = s->hmeta->nstars;
        // fixed-order sum on the host (ns doubles)
        double* h = nullptr;
        HIPCHK(hipHostMalloc((void**)&h, std::max(ns, 1LL) * 8, hipHostMallocDefault));
        hipError_t e = hipMemcpy(h, s->partial.p, (size_t)ns * 8, hipMemcpyDeviceToHost);
        double sum = 0;
        for (long long k = 0; k < ns; k++) sum += h[k];
        hipHostFree(h);
        if (e != hipSuccess) return set_err(GX_ERR_HIP, "copy of potential partials: %s", hipGetErrorString(e));
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
    if (s->pending) finish_step(s);
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

double gx_last_seconds(const gx_solver* s) { return s ? s->lastSeconds : 0; }

int gx_get_stats(const gx_solver* s, gx_stats* out) {
    if (!s || !out) return set_err(GX_ERR_ARG, "null argument");
    *out = s->stats;
    out->walk = s->walk;
    return GX_OK;
}

int64_t gx_export(gx_solver* s, int what, void* host, int64_t count) {
    if (!s || !host || count < 0) return set_err(GX_ERR_ARG, "bad argument"), -1;
    if (s->pending) finish_step(s);
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
