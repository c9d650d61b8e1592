// gx_sph.hip -- SPH hydro on the GPU: the neighbour BVH of internal/sph/tree.go
// and sph.Solver.Forces (internal/sph/sph.go:83-263), FP64.
//
// Pipeline of one evaluation (one stream, three host synchronisations: after
// the gas list, after the tree levels, at the end):
//   gas list + bounding box  -> keys over the box (tree.go:91-109)
//   stable radix sort (the gravity sort kernels) -> sorted copies (x, y, z, h)
//   tree levels: one launch per level, each node split where its keys first
//   differ (tree.go:126-161); children are numbered per level by atomics,
//   but the walk order does not depend on numbering: the walk visits right
//   children first, so leaves come by decreasing range start, the order of
//   the reference's stack walks (tree.go:205-267)
//   refit bottom up (boxes, hmax), density passes, refit, leaf sort by h
//   (tree.go:314-331), per-body terms, force pass.
//
// Walks: one wavefront takes 64 consecutive tree positions.  It walks the
// tree once for all of them, four nodes per step (a node is entered if any
// lane reaches it, the reference's own box test per lane); every lane keeps
// the leaves it accepts, tests their slots and appends its neighbours to its
// list in LDS; the lanes then sum their lists, each at full width.  Lists
// that fill up are worked off and the walk resumes from its stack.  So each
// body adds exactly the reference's neighbours, in its order: leaves by
// decreasing range start, bodies in leaf order.  Everything is FP64 with
// -ffp-contract=off; W, DW and the force terms keep the reference's
// expression order, and the smoothing-length target uses this project's
// correctly rounded cube root (the oracle uses the same sequence of IEEE
// operations).
#include <string.h>

#include <algorithm>
#include <cmath>

#include "gravitas_hip.h"
#include "gx_common.h"
#include "gx_kernels.h"

namespace gx {
namespace {

constexpr int kBvhLeaf = 8;        // leafSize, tree.go:17
constexpr int kBvhLevels = 256;    // tree levels the meta record holds
constexpr int kHyList = 24;        // neighbour slots per lane before they are summed
constexpr int kLeafList = 12;      // accepted leaves per lane before their slots are tested
constexpr int kDensityPasses = 6;  // maxDensityIterations, sph.go:41
constexpr double kPi = 3.14159265358979323846;

struct BvhNode {
    double minx, miny, minz, maxx, maxy, maxz, hmax;
    int lo, hi;  // range of tree positions
    int first;   // right child, walked first; -1 for a leaf
    int second;  // left child; -1 for a leaf
    int escape;  // next node after this subtree in a right-first walk; -1 ends it (exported)
    int parent;  // -1 for the root
};
static_assert(sizeof(BvhNode) == 80, "node record");

// What a walk reads per step, 256 bytes (one batch of scalar loads): the
// boxes and hmax of up to four nodes, the grandchildren of an inner node in
// the reference's walk order (right child's right, right child's left, left
// child's right, left child's left; a child that is a leaf stands for
// itself).  ref >= 0: an inner node, whose own record is quad[ref]; ref < 0:
// a leaf, ~(lo << 4 | count).
struct QuadNode {
    double minx[4], miny[4], minz[4], maxx[4], maxy[4], maxz[4], hmax[4];
    int ref[4];
    int count;
    int pad[3];
};
static_assert(sizeof(QuadNode) == 256, "walk record");

struct NodeBox {
    double minx, miny, minz, maxx, maxy, maxz, hmax;
};

struct HydroMeta {
    unsigned long long box[6];  // encoded min x, y, z and max x, y, z of the gas
    unsigned long long missing;  // gas bodies with h <= 0
    long long ngas;
    int lvlStart[kBvhLevels + 1];
    int lvlCount[kBvhLevels + 1];
    int leaves;
    int flags;
    unsigned long long evals, pairs, walkSteps;
    unsigned long long densityPairs;  // neighbour terms summed by the density passes
    unsigned long long viscPairs;     // force pairs that approach (viscosity branch)
    int maxPasses;
    int hChanged;  // some density pass moved a smoothing length (refitH is needed)
};

// Per tree position: the bodies' terms of the force pass.
struct HyTerms {
    double rho, c, f, sig4;  // density, sound speed, P / rho^2, 1 / (pi h^4)
};

// ---------------------------------------------------------------------------
// arithmetic in the reference's order

// Go's math.Min / math.Max and builtin min / max: NaN propagates, -0 < +0.
__host__ __device__ inline double go_min(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a == 0 && b == 0) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
__host__ __device__ inline double go_max(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a == 0 && b == 0) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}

// Cube root: reduction to m in [0.5, 4), Newton steps from a linear start,
// and a last Newton step on the residual y^3 - m computed exactly with FMA
// splits; correctly rounded in practice.  The oracle's or_cbrt is the same
// sequence of IEEE operations.
__host__ __device__ inline double hy_cbrt(double x) {
    if (!(x != 0) || std::isinf(x) || std::isnan(x)) return x;
    const bool neg = x < 0;
    const double a = neg ? -x : x;
    int e;
    const double f = std::frexp(a, &e);
    const int q = e >= 0 ? e / 3 : -((-e + 2) / 3);
    const int r = e - 3 * q;
    const double m = std::ldexp(f, r);
    double y = 0.6 + 0.25 * m;
    for (int it = 0; it < 5; it++) y = y - (y * y * y - m) / (3 * (y * y));
    const double t = y * y, tl = std::fma(y, y, -t);
    const double p = t * y, pl = std::fma(t, y, -p);
    const double res = (p - m) + (pl + tl * y);
    y = y - res / (3 * (y * y));
    y = std::ldexp(y, q);
    return neg ? -y : y;
}

// W with sigma = 1 / (pi h^3) of the same h (kernel.go:9-22).
__device__ inline double kernel_w(double r, double h, double sigma) {
    const double q = r / h;
    if (q < 1) return sigma * (1 - 1.5 * q * q + 0.75 * q * q * q);
    if (q < 2) {
        const double t = 2 - q;
        return sigma * 0.25 * t * t * t;
    }
    return 0;
}

// DW with sig4 = 1 / (pi h^4) of the same h (kernel.go:24-36).
__device__ inline double kernel_dw(double r, double h, double sig4) {
    const double q = r / h;
    if (q < 1) return sig4 * (-3 * q + 2.25 * q * q);
    if (q < 2) {
        const double t = 2 - q;
        return -sig4 * 0.75 * t * t;
    }
    return 0;
}

// boxDist2, tree.go:349-355, for pruning only.  A node is skipped when this
// distance reaches the lane's reach; every body j in the node's box has
// r2(j) >= box distance^2 in floating point too (rounding is monotone), so a
// skipped node holds no neighbour and the sums do not depend on how exactly
// this is rounded or on NaN/-0 handling: plain fmax, branch-free.
__device__ inline double box_dist2(const NodeBox& nd, double px, double py, double pz) {
    const double dx = fmax(fmax(nd.minx - px, px - nd.maxx), 0.0);
    const double dy = fmax(fmax(nd.miny - py, py - nd.maxy), 0.0);
    const double dz = fmax(fmax(nd.minz - pz, pz - nd.maxz), 0.0);
    return dx * dx + dy * dy + dz * dz;
}

__host__ inline double dec_host(unsigned long long u) {
    const unsigned long long v = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    double d;
    memcpy(&d, &v, sizeof d);
    return d;
}

// Wavefront of block b when consecutive waves share an XCD: the grid is
// 8 * ceil(nw / 8) blocks and block b runs wave (b % 8) * per + b / 8.
__device__ inline long long xcd_wave(long long nw) {
    const long long per = (nw + 7) / 8;
    return (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

__device__ inline unsigned long long wave_min_u64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long o = __shfl_xor(v, m, kWave);
        v = o < v ? o : v;
    }
    return v;
}
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long o = __shfl_xor(v, m, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// ---------------------------------------------------------------------------
// gas list, box, keys

__global__ void k_hy_meta_init(HydroMeta* meta) {
    const int t = threadIdx.x;
    if (t < 3) meta->box[t] = ~0ull;
    if (t >= 3 && t < 6) meta->box[t] = 0;
    if (t == 0) {
        meta->missing = 0;
        meta->ngas = 0;
        meta->leaves = 0;
        meta->flags = 0;
        meta->evals = meta->pairs = meta->walkSteps = 0;
        meta->densityPairs = meta->viscPairs = 0;
        meta->maxPasses = 0;
        meta->hChanged = 0;
    }
    for (int l = t; l <= kBvhLevels; l += blockDim.x) {
        meta->lvlStart[l] = 0;
        meta->lvlCount[l] = 0;
    }
}

// Gas flag, bounding box of the gas and the count of missing smoothing
// lengths (tree.go:93-98, sph.go:279-288).
__global__ __launch_bounds__(256) void k_hy_flag(long long n, const uint8_t* __restrict__ kind,
                                                 const uint8_t* __restrict__ alive, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ h, uint8_t* __restrict__ flag,
                                                 HydroMeta* meta) {
    __shared__ unsigned long long sb[4][7];
    unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0}, miss = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const bool g = kind[i] == kGas && alive[i];
        flag[i] = g;
        if (!g) continue;
        const double p[3] = {x[i], y[i], z[i]};
        for (int c = 0; c < 3; c++) {
            const unsigned long long e = enc_d(p[c]);
            mn[c] = e < mn[c] ? e : mn[c];
            mx[c] = e > mx[c] ? e : mx[c];
        }
        miss += h[i] <= 0;
    }
    const int wv = threadIdx.x / kWave;
    for (int c = 0; c < 3; c++) {
        mn[c] = wave_min_u64(mn[c]);
        mx[c] = wave_max_u64(mx[c]);
    }
    miss = wave_sum_u64(miss);
    if (lane_id() == 0) {
        for (int c = 0; c < 3; c++) {
            sb[wv][c] = mn[c];
            sb[wv][3 + c] = mx[c];
        }
        sb[wv][6] = miss;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int c = threadIdx.x;
        unsigned long long v = sb[0][c];
        for (int w = 1; w < (int)(blockDim.x / kWave); w++) {
            const unsigned long long o = sb[w][c];
            v = c < 3 ? (o < v ? o : v) : c < 6 ? (o > v ? o : v) : v + o;
        }
        if (c < 3) atomicMin(&meta->box[c], v);
        else if (c < 6) atomicMax(&meta->box[c], v);
        else if (v) atomicAdd(&meta->missing, v);
    }
}

// initH: the first guess for the smoothing lengths that are missing.
__global__ void k_hy_init_h(const int32_t* __restrict__ gas, const HydroMeta* meta, double guess,
                            double* __restrict__ h) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= meta->ngas) return;
    const int32_t i = gas[k];
    if (h[i] <= 0) h[i] = guess;
}

// Keys over the gas box, tree.go:99-109; the value is the body index.
__global__ void k_hy_keys(const int32_t* __restrict__ gas, const HydroMeta* meta, const double* __restrict__ x,
                          const double* __restrict__ y, const double* __restrict__ z, double minX, double minY,
                          double minZ, double scale, uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= meta->ngas) return;
    const int32_t i = gas[k];
    const uint64_t kx = spread21((uint64_t)((x[i] - minX) * scale));
    const uint64_t ky = spread21((uint64_t)((y[i] - minY) * scale));
    const uint64_t kz = spread21((uint64_t)((z[i] - minZ) * scale));
    keys[k] = kx << 2 | ky << 1 | kz;
    vals[k] = i;
}

__global__ void k_hy_sort_meta(const HydroMeta* hm, Meta* meta) { meta->nstars = hm->ngas; }

// Sorted copies (tree.go:113-119) and masses.
__global__ void k_hy_gather(const int32_t* __restrict__ order, const HydroMeta* meta, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z,
                            const double* __restrict__ h, const double* __restrict__ m, double4* __restrict__ sp,
                            double* __restrict__ sm) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= meta->ngas) return;
    const int32_t i = order[k];
    sp[k] = make_double4(x[i], y[i], z[i], h[i]);
    sm[k] = m[i];
}

// ---------------------------------------------------------------------------
// tree

__global__ void k_hy_root(HydroMeta* meta, BvhNode* nodes) {
    BvhNode r = {};
    r.lo = 0;
    r.hi = (int)meta->ngas;
    r.first = r.second = -1;
    r.escape = -1;
    r.parent = -1;
    nodes[0] = r;
    meta->lvlStart[0] = 0;
    meta->lvlCount[0] = 1;
}

// split, tree.go:150-161.
__device__ inline int split_range(const uint64_t* __restrict__ keys, int lo, int hi) {
    const uint64_t a = keys[lo], b = keys[hi - 1];
    if (a == b) return (lo + hi) / 2;
    const uint64_t bit = 1ull << (63 - __clzll((long long)(a ^ b)));
    int l = 0, r = hi - lo;
    while (l < r) {
        const int mid = l + (r - l) / 2;
        if (keys[lo + mid] & bit) r = mid;
        else l = mid + 1;
    }
    return lo + l;
}

// Tree level L: leaves (<= 8 bodies) are marked, the other nodes split into
// a left and a right child on level L + 1 (buildNode, tree.go:126-146).
__global__ __launch_bounds__(256) void k_hy_level(int L, const uint64_t* __restrict__ keys,
                                                  BvhNode* __restrict__ nodes, HydroMeta* meta,
                                                  int2* __restrict__ leafRange) {
    const int start = meta->lvlStart[L], cnt = meta->lvlCount[L];
    const int next = start + cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) meta->lvlStart[L + 1] = next;
    int leaves = 0;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < cnt; idx += gridDim.x * blockDim.x) {
        BvhNode* nd = &nodes[start + idx];
        const int lo = nd->lo, hi = nd->hi;
        if (hi - lo <= kBvhLeaf) {
            nd->first = nd->second = -1;
            for (int k = lo; k < hi; k++) leafRange[k] = make_int2(lo, hi);
            leaves++;
            continue;
        }
        const int mid = split_range(keys, lo, hi);
        const int left = next + atomicAdd(&meta->lvlCount[L + 1], 2), right = left + 1;
        BvhNode c = {};
        c.first = c.second = -1;
        c.parent = start + idx;
        c.lo = lo;
        c.hi = mid;
        c.escape = nd->escape;
        nodes[left] = c;
        c.lo = mid;
        c.hi = hi;
        c.escape = left;
        nodes[right] = c;
        nd->first = right;
        nd->second = left;
    }
    leaves = wave_sum_i(leaves);
    if (lane_id() == 0 && leaves) atomicAdd(&meta->leaves, leaves);
}

// Boxes and hmax of the nodes of one level from the sorted copies (leaves;
// hmax from h through the order when h is given) or from the children
// (refit, tree.go:164-180; refitH, :184-201).
__device__ inline void refit_node(BvhNode* __restrict__ nodes, int ni,
                                  const double4* __restrict__ sp, const int32_t* __restrict__ order,
                                  const double* __restrict__ h) {
    BvhNode* nd = &nodes[ni];
    double mnx = INFINITY, mny = INFINITY, mnz = INFINITY, mxx = -INFINITY, mxy = -INFINITY, mxz = -INFINITY;
    double hm = 0;
    if (nd->first < 0) {
        for (int k = nd->lo; k < nd->hi; k++) {
            const double4 p = sp[k];
            mnx = go_min(mnx, p.x); mxx = go_max(mxx, p.x);
            mny = go_min(mny, p.y); mxy = go_max(mxy, p.y);
            mnz = go_min(mnz, p.z); mxz = go_max(mxz, p.z);
            hm = go_max(hm, h ? h[order[k]] : p.w);
        }
    } else {
        const BvhNode* c[2] = {&nodes[nd->second], &nodes[nd->first]};  // left, right
        for (int s = 0; s < 2; s++) {
            mnx = go_min(mnx, c[s]->minx); mxx = go_max(mxx, c[s]->maxx);
            mny = go_min(mny, c[s]->miny); mxy = go_max(mxy, c[s]->maxy);
            mnz = go_min(mnz, c[s]->minz); mxz = go_max(mxz, c[s]->maxz);
            hm = go_max(hm, c[s]->hmax);
        }
    }
    nd->minx = mnx; nd->miny = mny; nd->minz = mnz;
    nd->maxx = mxx; nd->maxy = mxy; nd->maxz = mxz;
    nd->hmax = hm;
}

// One wide level, a grid over its nodes.
__global__ __launch_bounds__(256) void k_hy_refit(int L, const double4* __restrict__ sp,
                                                  const int32_t* __restrict__ order, const double* __restrict__ h,
                                                  BvhNode* __restrict__ nodes, const HydroMeta* meta) {
    if (h && !meta->hChanged) return;  // refitH only when a smoothing length moved (sph.go:178-180)
    const int start = meta->lvlStart[L], cnt = meta->lvlCount[L];
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < cnt; idx += gridDim.x * blockDim.x)
        refit_node(nodes, start + idx, sp, order, h);
}

// Levels Lhi down to Llo, all small, in one workgroup: the waves of a
// workgroup share one CU and its L1, so a barrier between levels is enough
// for the children to be seen.
__global__ __launch_bounds__(1024) void k_hy_refit_small(int Lhi, int Llo, const double4* __restrict__ sp,
                                                         const int32_t* __restrict__ order,
                                                         const double* __restrict__ h, BvhNode* __restrict__ nodes,
                                                         const HydroMeta* meta) {
    if (h && !meta->hChanged) return;
    for (int L = Lhi; L >= Llo; L--) {
        const int start = meta->lvlStart[L], cnt = meta->lvlCount[L];
        for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) refit_node(nodes, start + idx, sp, order, h);
        __syncthreads();
    }
}

// The walk record of every inner node (and of the root when it is a leaf):
// its grandchildren in walk order, with their boxes and hmax.
__global__ __launch_bounds__(256) void k_hy_quad(const BvhNode* __restrict__ nodes, long long nnodes,
                                                 QuadNode* __restrict__ quad, const int* __restrict__ onlyIf) {
    const long long ni = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ni >= nnodes || (onlyIf && !*onlyIf)) return;
    const BvhNode& nd = nodes[ni];
    int f[4], n = 0;
    if (nd.first < 0) {
        if (ni != 0) return;
        f[n++] = 0;  // a root leaf: a walk of one leaf
    } else {
        const int c[2] = {nd.first, nd.second};  // right, left
        for (int s = 0; s < 2; s++) {
            const BvhNode& cn = nodes[c[s]];
            if (cn.first >= 0) {
                f[n++] = cn.first;
                f[n++] = cn.second;
            } else {
                f[n++] = c[s];
            }
        }
    }
    QuadNode q;
    for (int s = 0; s < 4; s++) {
        if (s < n) {
            const BvhNode& e = nodes[f[s]];
            q.minx[s] = e.minx; q.miny[s] = e.miny; q.minz[s] = e.minz;
            q.maxx[s] = e.maxx; q.maxy[s] = e.maxy; q.maxz[s] = e.maxz;
            q.hmax[s] = e.hmax;
            q.ref[s] = e.first >= 0 ? f[s] : ~(e.lo << 4 | (e.hi - e.lo));
        } else {
            q.minx[s] = q.miny[s] = q.minz[s] = INFINITY;
            q.maxx[s] = q.maxy[s] = q.maxz[s] = -INFINITY;
            q.hmax[s] = 0;
            q.ref[s] = 0;
        }
    }
    q.count = n;
    q.pad[0] = q.pad[1] = q.pad[2] = 0;
    quad[ni] = q;
}

// sortLeaves, tree.go:314-331: in every leaf the bodies by decreasing h
// (the final h, read through the order), stable; out of place, one thread per
// position; the sorted copies carry the final h.
__global__ void k_hy_sort_leaves(const HydroMeta* meta, const int2* __restrict__ leafRange,
                                 const int32_t* __restrict__ order, const double4* __restrict__ sp,
                                 const double* __restrict__ h, int32_t* __restrict__ order2,
                                 double4* __restrict__ sp2) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= meta->ngas) return;
    const int2 lr = leafRange[k];
    const int32_t i = order[k];
    const double hk = h[i];
    int rank = 0;
    for (int f = lr.x; f < lr.y; f++) {
        const double hf = h[order[f]];
        rank += hf > hk || (hf == hk && f < k);
    }
    order2[lr.x + rank] = i;
    const double4 p = sp[k];
    sp2[lr.x + rank] = make_double4(p.x, p.y, p.z, hk);
}

// ---------------------------------------------------------------------------
// walks

// The walk over the tree for the 64 bodies of a wavefront, resumable: the
// stack (LDS, wave-uniform entries: ref, and for a leaf the mask of lanes
// that accept it) holds what is still to visit, top first.  Popping an inner
// node loads its record and tests its (up to four) grandchildren per lane
// (visit); those some lane accepts are pushed in reverse, so they pop in walk
// order.  Popping a leaf calls leaf(lo, count, in) (in: this lane accepts
// it); leaf returns true when a lane's list is full, and the walk returns
// with the stack as it stands.  A descendant is accepted only where its
// ancestors are (smaller box, smaller hmax), so skipping the tests of the
// children in between loses no leaf.
struct WalkStack {
    int* ref;
    unsigned long long* mask;
    int cap;
    int top;
};

template <class Visit, class Leaf>
__device__ inline void walk_tree(const QuadNode* __restrict__ quad, WalkStack& s, Visit visit, Leaf leaf,
                                 unsigned long long& steps, int* overflow) {
    const int lane = lane_id();
    while (s.top > 0) {
        s.top--;
        const int ref = __builtin_amdgcn_readfirstlane(s.ref[s.top]);
        if (ref < 0) {
            const unsigned long long m = s.mask[s.top];
            const int e = ~ref;
            if (leaf(e >> 4, e & 15, ((m >> lane) & 1) != 0)) return;
            continue;
        }
        steps++;
        const QuadNode& q = quad[ref];
        const int n = q.count;
        unsigned long long m[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const NodeBox b = {q.minx[c], q.miny[c], q.minz[c], q.maxx[c], q.maxy[c], q.maxz[c], q.hmax[c]};
            m[c] = __ballot(visit(b)) & (c < n ? ~0ull : 0ull);
        }
#pragma unroll
        for (int c = 3; c >= 0; c--) {
            if (m[c] == 0) continue;
            if (s.top >= s.cap) {  // cannot happen with the host's capacity; abandon the walk, report it
                if (lane == 0) atomicOr(overflow, 1);
                s.top = 0;
                return;
            }
            if (lane == 0) {
                s.ref[s.top] = q.ref[c];
                s.mask[s.top] = m[c];
            }
            s.top++;
        }
    }
}

__device__ inline int wave_max_i(int v) {
    for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, kWave));
    return v;
}

// Neighbours of the wavefront's bodies in the reference's order.  The walk is
// shared; every lane keeps the leaves it accepts (leaf list, LDS, kLeafList
// per lane), then tests the slots of its own leaves (slot(q, k) on the sorted
// copy q of tree position k) and appends the hits to its neighbour list (LDS,
// kHyList per lane); sum(list, cnt) works off the neighbour lists, in order.
// Leaves come in walk order (decreasing range start), slots in leaf order, so
// each body sees its neighbours in the order of the reference's scans.
template <class Visit, class Slot, class Sum>
__device__ inline void neighbour_walk(const QuadNode* __restrict__ quad, const double4* __restrict__ sp,
                                      int* __restrict__ llist, int* __restrict__ nlist, WalkStack& stk, int lane,
                                      Visit visit, Slot slot, Sum sum, unsigned long long& steps, int* overflow) {
    int nl = 0, cnt = 0;
    stk.top = 1;
    if (lane == 0) {
        stk.ref[0] = 0;
        stk.mask[0] = ~0ull;
    }
    while (stk.top > 0) {
        walk_tree(
            quad, stk, visit,
            [&](int lo, int count, bool in) {
                if (in) llist[nl++ * kWave + lane] = lo << 4 | count;
                return __ballot(nl >= kLeafList) != 0;
            },
            steps, overflow);
        const int maxl = wave_max_i(nl);
        for (int e = 0; e < maxl; e++) {
            const int ent = e < nl ? llist[e * kWave + lane] : 0;
            const int lo = ent >> 4, count = ent & 15;
            const int last = max(count, 1) - 1;
#pragma unroll
            for (int j0 = 0; j0 < kBvhLeaf; j0 += 4) {  // the slots in two batches of gathers
                double4 q[4];
#pragma unroll
                for (int j = 0; j < 4; j++) q[j] = sp[lo + min(j0 + j, last)];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (j0 + j < count && slot(q[j], lo + j0 + j)) nlist[cnt++ * kWave + lane] = lo + j0 + j;
            }
            if (__ballot(cnt > kHyList - kBvhLeaf) != 0) {
                sum(nlist, cnt);
                cnt = 0;
            }
        }
        nl = 0;
    }
    sum(nlist, cnt);
}

// LDS of a walk kernel: leaf lists, neighbour lists, then the stack.
__host__ __device__ inline size_t walk_lds_bytes(int stackCap) {
    return sizeof(int) * (kLeafList + kHyList) * kWave + (sizeof(int) + sizeof(unsigned long long)) * stackCap + 16;
}

// Density passes and smoothing lengths (sph.go:115-182).  Lane = tree
// position; each evaluated body iterates on its own h, at most 6 passes.
extern __shared__ __align__(16) unsigned char hy_lds[];

// The LDS of a walk kernel carved into leaf lists, neighbour lists and stack.
struct WalkLds {
    int* llist;
    int* nlist;
    WalkStack stk;
};

__device__ inline WalkLds walk_lds(int stackCap) {
    WalkLds l;
    l.llist = (int*)hy_lds;
    l.nlist = l.llist + kLeafList * kWave;
    unsigned long long* mask = (unsigned long long*)(l.nlist + kHyList * kWave);  // 8-byte aligned
    l.stk.mask = mask;
    l.stk.ref = (int*)(mask + stackCap);
    l.stk.cap = stackCap;
    l.stk.top = 0;
    return l;
}

__global__ __launch_bounds__(64) void k_hy_density(const QuadNode* __restrict__ quad, int stackCap,
                                                   const double4* __restrict__ sp,
                                                   const double* __restrict__ sm, const int32_t* __restrict__ order,
                                                   const uint8_t* __restrict__ active, HydroMeta* meta, double hMin,
                                                   double ngb, double* __restrict__ h, double* __restrict__ rho) {
    WalkLds L = walk_lds(stackCap);
    const long long ngas = meta->ngas;
    const long long nw = (ngas + kWave - 1) / kWave;
    const long long w = xcd_wave(nw);
    if (w >= nw) return;
    const int lane = lane_id();
    const long long k = w * kWave + lane;
    const bool valid = k < ngas;
    const int32_t i = valid ? order[k] : 0;
    const bool act = valid && (active == nullptr || active[i]);
    const double4 p = valid ? sp[k] : make_double4(0, 0, 0, 1);
    const double mi = valid ? sm[k] : 0;
    double hi = p.w, rhoi = 0;
    bool need = act;
    int passes = 0;
    unsigned long long walkSteps = 0, dpairs = 0;
    for (int it = 0; it < kDensityPasses; it++) {
        if (__ballot(need) == 0) break;
        const double radius = 2 * hi;
        const double rr = radius * radius;
        const double sigma = 1 / (kPi * hi * hi * hi);
        double rsum = 0;
        neighbour_walk(
            quad, sp, L.llist, L.nlist, L.stk, lane,
            [&](const NodeBox& nd) { return !(box_dist2(nd, p.x, p.y, p.z) >= rr) & need; },
            [&](const double4& q, int) {
                const double ex = q.x - p.x, ey = q.y - p.y, ez = q.z - p.z;
                return need && ex * ex + ey * ey + ez * ez < rr;
            },
            [&](const int* list, int cnt) {
                for (int e = 0; e < cnt; e++) {
                    const int kk = list[e * kWave + lane];
                    const double4 q = sp[kk];
                    const double ex = q.x - p.x, ey = q.y - p.y, ez = q.z - p.z;
                    const double r2 = ex * ex + ey * ey + ez * ez;
                    rsum += sm[kk] * kernel_w(sqrt(r2), hi, sigma);
                }
                dpairs += cnt;
            },
            walkSteps, &meta->flags);
        if (need) {  // setH, sph.go:121-131
            passes++;
            rhoi = rsum;
            const double target = go_max(hMin, 0.5 * hy_cbrt(3 * ngb * mi / (4 * kPi * rsum)));
            const bool last = it == kDensityPasses - 1;
            const bool redo = !last && fabs(target - hi) > 0.05 * hi;
            if (redo) hi = go_max(hMin, go_max(0.5 * hi, go_min(2 * hi, target)));
            need = redo;
        }
    }
    if (act) {
        rho[i] = rhoi;
        h[i] = hi;
    }
    if (__ballot(act && hi != p.w) != 0 && lane == 0) atomicOr(&meta->hChanged, 1);
    const unsigned long long ev = wave_sum_u64((unsigned long long)passes);
    const unsigned long long dp = wave_sum_u64(dpairs);
    const int mp = wave_max_i(passes);
    if (lane == 0) {
        if (ev) atomicAdd(&meta->evals, ev);
        if (dp) atomicAdd(&meta->densityPairs, dp);
        atomicMax(&meta->maxPasses, mp);
    }
}

// Per tree position (after the leaf sort): velocities and mass, and the
// terms P / rho^2, sound speed and 1 / (pi h^4) (sph.go:186-194).
__global__ void k_hy_terms(const int32_t* __restrict__ order, const HydroMeta* meta, const double4* __restrict__ sp,
                           const double* __restrict__ vx, const double* __restrict__ vy,
                           const double* __restrict__ vz, const double* __restrict__ m,
                           const double* __restrict__ rho, const double* __restrict__ u,
                           const double* __restrict__ step, double gamma, double4* __restrict__ sv,
                           HyTerms* __restrict__ st, double* __restrict__ sstep) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= meta->ngas) return;
    const int32_t i = order[k];
    sv[k] = make_double4(vx[i], vy[i], vz[i], m[i]);
    const double ui = go_max(u[i], 0.0);
    const double ri = rho[i];
    const double P = (gamma - 1) * ri * ui;
    const double h = sp[k].w;
    HyTerms t;
    t.rho = ri;
    t.c = sqrt(gamma * (gamma - 1) * ui);
    t.f = P / (ri * ri);
    t.sig4 = 1 / (kPi * h * h * h * h);
    st[k] = t;
    if (step) sstep[k] = step[i];
}

// lowerWake, sph.go:266-274.
__device__ inline void lower_wake(unsigned long long* w, double dt) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(dt);
    unsigned long long old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if (__longlong_as_double((long long)old) <= dt) return;
        const unsigned long long prev = atomicCAS(w, old, bits);
        if (prev == old) return;
        old = prev;
    }
}

// Forces, energy equation and velocity divergence of the evaluated bodies
// (sph.go:196-261), and the wake of neighbours with too long steps.
__global__ __launch_bounds__(64) void k_hy_force(const QuadNode* __restrict__ quad, int stackCap,
                                                 const double4* __restrict__ sp,
                                                 const double4* __restrict__ sv, const HyTerms* __restrict__ st,
                                                 const double* __restrict__ sstep, const int32_t* __restrict__ order,
                                                 const uint8_t* __restrict__ active, HydroMeta* meta, double alpha,
                                                 double* __restrict__ ax, double* __restrict__ ay,
                                                 double* __restrict__ az, double* __restrict__ dudt,
                                                 double* __restrict__ divv, double* __restrict__ dtOut,
                                                 unsigned long long* __restrict__ wake) {
    WalkLds L = walk_lds(stackCap);
    const long long ngas = meta->ngas;
    const long long nw = (ngas + kWave - 1) / kWave;
    const long long w = xcd_wave(nw);
    if (w >= nw) return;
    const int lane = lane_id();
    const long long k = w * kWave + lane;
    const bool valid = k < ngas;
    const int32_t i = valid ? order[k] : 0;
    const bool act = valid && (active == nullptr || active[i]);
    const double4 p = valid ? sp[k] : make_double4(0, 0, 0, 1);
    const double4 v = valid ? sv[k] : make_double4(0, 0, 0, 0);
    const HyTerms ti = valid ? st[k] : HyTerms{1, 0, 0, 0};
    const double hi = p.w, rhoi = ti.rho, ci = ti.c, fi = ti.f;
    const double radius = 2 * hi;
    const double r2max = radius * radius;
    double fx = 0, fy = 0, fz = 0, du = 0, dv = 0;
    double vsig = 2 * ci, minStep = INFINITY, maxStep = 0.0;
    unsigned long long walkSteps = 0, pairs = 0, viscPairs = 0;
    // scan's test of a body slot (tree.go:295-307)
    auto reaches = [&](const double4& q, double& r2) {
        const double ex = q.x - p.x, ey = q.y - p.y, ez = q.z - p.z;
        r2 = ex * ex + ey * ey + ez * ez;
        double lim2 = r2max;
        const double l = 2 * q.w;
        if (l * l > lim2) lim2 = l * l;
        return r2 < lim2;
    };
    // each's node test with symmetric set (tree.go:214-221)
    auto enters = [&](const NodeBox& nd, bool on) {
        double reach = radius;
        if (2 * nd.hmax > reach) reach = 2 * nd.hmax;
        return !(box_dist2(nd, p.x, p.y, p.z) >= reach * reach) & on;
    };
    neighbour_walk(
        quad, sp, L.llist, L.nlist, L.stk, lane, [&](const NodeBox& nd) { return enters(nd, act); },
        [&](const double4& q, int kk) {
            double r2;
            return act && reaches(q, r2) && kk != k && r2 != 0;
        },
        [&](const int* list, int cnt) {
            for (int e = 0; e < cnt; e++) {
                const int kk = list[e * kWave + lane];
                const double4 q = sp[kk];
                const double4 sj = sv[kk];
                const HyTerms tj = st[kk];
                const double ex0 = q.x - p.x, ey0 = q.y - p.y, ez0 = q.z - p.z;
                const double r2 = ex0 * ex0 + ey0 * ey0 + ez0 * ez0;
                if (sstep) {
                    const double sj0 = sstep[kk];
                    if (sj0 > 0) {
                        minStep = go_min(minStep, sj0);
                        maxStep = go_max(maxStep, sj0);
                    }
                }
                const double r = sqrt(r2);
                const double ex = (p.x - q.x) / r, ey = (p.y - q.y) / r, ez = (p.z - q.z) / r;
                const double vx = v.x - sj.x, vy = v.y - sj.y, vz = v.z - sj.z;
                const double vr = vx * ex + vy * ey + vz * ez;
                vsig = go_max(vsig, ci + tj.c - 3 * go_min(0.0, vr));
                const double dwi = kernel_dw(r, hi, ti.sig4), dwj = kernel_dw(r, q.w, tj.sig4);
                const double dwm = 0.5 * (dwi + dwj);
                const double fj = tj.f;
                double visc = 0;
                if (vr < 0) {
                    viscPairs++;
                    const double hm = 0.5 * (hi + q.w);
                    const double mu = hm * vr * r / (r2 + 0.01 * hm * hm);
                    visc = (-alpha * 0.5 * (ci + tj.c) * mu + 2 * alpha * mu * mu) / (0.5 * (rhoi + tj.rho));
                }
                const double f = -sj.w * (fi * dwi + fj * dwj + visc * dwm);
                fx = fx + f * ex;
                fy = fy + f * ey;
                fz = fz + f * ez;
                du += sj.w * (fi * dwi + 0.5 * visc * dwm) * vr;
                dv -= sj.w * dwi * vr;
            }
            pairs += cnt;
        },
        walkSteps, &meta->flags);
    double dt = 4 * minStep;  // LimiterRatio, internal/sim/sim.go:90
    if (vsig > 0) dt = go_min(dt, 0.3 * hi / vsig);
    if (act) {
        ax[i] = fx;
        ay[i] = fy;
        az[i] = fz;
        dudt[i] = du;
        divv[i] = dv / rhoi;
        dtOut[i] = dt;
    }
    // Neighbours with much longer steps must be woken up (sph.go:250-258);
    // j = i and r2 = 0 count here, as in the reference's second scan.
    const bool wakeNeeded = act && sstep != nullptr && maxStep > 4 * dt;
    if (__ballot(wakeNeeded) != 0) {
        L.stk.top = 1;
        if (lane == 0) {
            L.stk.ref[0] = 0;
            L.stk.mask[0] = ~0ull;
        }
        walk_tree(
            quad, L.stk, [&](const NodeBox& nd) { return enters(nd, wakeNeeded); },
            [&](int lo, int count, bool) {
                double4 q[kBvhLeaf];
#pragma unroll
                for (int j = 0; j < kBvhLeaf; j++) q[j] = sp[lo + min(j, count - 1)];
#pragma unroll
                for (int j = 0; j < kBvhLeaf; j++) {
                    if (j >= count) break;
                    const int kk = lo + j;
                    double r2;
                    if (wakeNeeded && reaches(q[j], r2) && sstep[kk] > 4 * dt) lower_wake(&wake[order[kk]], dt);
                }
                return false;
            },
            walkSteps, &meta->flags);
    }
    const unsigned long long pw = wave_sum_u64(pairs), vw = wave_sum_u64(viscPairs);
    if (lane == 0) {
        if (pw) atomicAdd(&meta->pairs, pw);
        if (vw) atomicAdd(&meta->viscPairs, vw);
        atomicAdd(&meta->walkSteps, walkSteps);
    }
}

// Fewer than two gas bodies (sph.go:91-100).
__global__ void k_hy_few(const int32_t* __restrict__ gas, const HydroMeta* meta, const uint8_t* __restrict__ active,
                         double* ax, double* ay, double* az, double* dudt, double* divv, double* dt) {
    const long long k = threadIdx.x;
    if (k >= meta->ngas) return;
    const int32_t i = gas[k];
    if (active && !active[i]) return;
    ax[i] = ay[i] = az[i] = 0;
    dudt[i] = divv[i] = 0;
    dt[i] = INFINITY;
}

int blocks_of(long long n, int t) { return (int)std::max<long long>(1, (n + t - 1) / t); }

#define HYCHK(x)                                                                                         \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return set_error(GX_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));        \
    } while (0)

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

int grow(Buf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return GX_OK;
    if (b.p) HYCHK(hipFree(b.p));
    b.p = nullptr;
    const size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    HYCHK(hipMalloc(&b.p, want));
    b.bytes = want;
    return GX_OK;
}

constexpr int kHyPhases = 5;

}  // namespace
}  // namespace gx

using namespace gx;

struct gx_hydro {
    hipStream_t own = nullptr, stream = nullptr;
    int timing = 0;
    Buf meta, sortMeta, flag, tileCnt, gas, keys, keys2, vals, vals2, hist, digitTotal;
    Buf sp, sp2, sm, order2, nodes, quad, leafRange, sv, terms, sstep;
    Buf stage[21];  // host-array entry point
    HydroMeta* hmeta = nullptr;  // pinned mirror
    int depth = 0;               // levels of the last tree, for the next build's first batch
    long long nnodes = 0, ngas = 0;
    hipEvent_t ev[kHyPhases + 1] = {};
    float phase[kHyPhases + 1] = {};
    gx_hydro_stats stats = {};
};

namespace {

int hydro_run(gx_hydro* H, const double* x, const double* y, const double* z, const double* vx, const double* vy,
              const double* vz, const double* m, const uint8_t* kind, const uint8_t* alive, const double* u,
              double* h, double* rho, long long n, const gx_gas_params* gp, double softening, const uint8_t* active,
              const double* step, double* ax, double* ay, double* az, double* dudt, double* divv, double* dtOut,
              uint64_t* wake) {
    hipStream_t st = H->stream;
    auto mark = [&](int p) {
        if (H->timing) (void)hipEventRecord(H->ev[p], st);
    };
    int rc;
    const size_t n1 = (size_t)std::max<long long>(n, 1);
    const long long ntilesFlag = (n + kCollectTile - 1) / kCollectTile;
    if ((rc = grow(H->meta, sizeof(HydroMeta))) || (rc = grow(H->sortMeta, sizeof(Meta))) ||
        (rc = grow(H->flag, n1)) || (rc = grow(H->tileCnt, sizeof(int) * std::max<long long>(ntilesFlag, 1))) ||
        (rc = grow(H->gas, sizeof(int32_t) * n1)))
        return rc;
    HydroMeta* meta = (HydroMeta*)H->meta.p;
    mark(0);
    hipLaunchKernelGGL(k_hy_meta_init, dim3(1), dim3(256), 0, st, meta);
    hipLaunchKernelGGL(k_hy_flag, dim3((unsigned)std::min<long long>(1024, blocks_of(n, 256))), dim3(256), 0, st, n,
                       kind, alive, x, y, z, (const double*)h, (uint8_t*)H->flag.p, meta);
    hipLaunchKernelGGL(k_flag_count, dim3((unsigned)std::max<long long>(ntilesFlag, 1)), dim3(256), 0, st, n,
                       (const uint8_t*)H->flag.p, (int*)H->tileCnt.p);
    hipLaunchKernelGGL(k_flag_scan, dim3(1), dim3(1024), 0, st, (int)std::max<long long>(ntilesFlag, 1),
                       (int*)H->tileCnt.p, &meta->ngas);
    hipLaunchKernelGGL(k_flag_scatter, dim3((unsigned)std::max<long long>(ntilesFlag, 1)), dim3(256), 0, st, n,
                       (const uint8_t*)H->flag.p, (const int*)H->tileCnt.p, (int32_t*)H->gas.p);
    HYCHK(hipMemcpyAsync(H->hmeta, meta, offsetof(HydroMeta, lvlStart), hipMemcpyDeviceToHost, st));
    HYCHK(hipStreamSynchronize(st));
    const long long ngas = H->hmeta->ngas;
    H->ngas = ngas;
    memset(&H->stats, 0, sizeof H->stats);
    memset(H->phase, 0, sizeof H->phase);  // an early return leaves no stale phase times
    H->stats.n_gas = ngas;
    const int32_t* gas = (const int32_t*)H->gas.p;
    if (ngas < 2) {
        if (ngas == 1)
            hipLaunchKernelGGL(k_hy_few, dim3(1), dim3(64), 0, st, gas, meta, active, ax, ay, az, dudt, divv, dtOut);
        H->nnodes = 0;
        HYCHK(hipGetLastError());
        HYCHK(hipStreamSynchronize(st));
        return GX_OK;
    }
    const double gamma = gp && gp->gamma != 0 ? gp->gamma : 5.0 / 3;
    const double ngb = gp && gp->neighbors != 0 ? gp->neighbors : 40;
    const double alpha = gp && gp->alpha != 0 ? gp->alpha : 1;
    const double minX = dec_host(H->hmeta->box[0]), minY = dec_host(H->hmeta->box[1]),
                 minZ = dec_host(H->hmeta->box[2]);
    const double maxX = dec_host(H->hmeta->box[3]), maxY = dec_host(H->hmeta->box[4]),
                 maxZ = dec_host(H->hmeta->box[5]);
    if (H->hmeta->missing) {  // initH, sph.go:277-302
        const double vol = go_max((maxX - minX) * (maxY - minY) * (maxZ - minZ), 1e-30);
        const double guess = 0.5 * hy_cbrt(3 * ngb * vol / (4 * kPi * (double)ngas));
        hipLaunchKernelGGL(k_hy_init_h, dim3(blocks_of(ngas, 256)), dim3(256), 0, st, gas, meta, guess, h);
    }
    const double side = go_max(go_max(maxX - minX, maxY - minY), go_max(maxZ - minZ, 1e-12)) * (1 + 1e-9);
    const double scale = (double)((1 << 21) - 1) / side;
    const size_t g1 = (size_t)ngas;
    const int ntiles = (int)std::max<long long>(1, (ngas + kSortTile - 1) / kSortTile);
    const size_t maxNodes = 2 * g1 + 1;
    if ((rc = grow(H->keys, 8 * g1)) || (rc = grow(H->keys2, 8 * g1)) || (rc = grow(H->vals, 4 * g1)) ||
        (rc = grow(H->vals2, 4 * g1)) || (rc = grow(H->hist, sizeof(uint32_t) * 256 * (size_t)ntiles)) ||
        (rc = grow(H->digitTotal, sizeof(uint32_t) * 256)) || (rc = grow(H->sp, sizeof(double4) * g1)) ||
        (rc = grow(H->sp2, sizeof(double4) * g1)) || (rc = grow(H->sm, 8 * g1)) ||
        (rc = grow(H->order2, 4 * g1)) || (rc = grow(H->nodes, sizeof(BvhNode) * maxNodes)) ||
        (rc = grow(H->quad, sizeof(QuadNode) * maxNodes)) ||
        (rc = grow(H->leafRange, sizeof(int2) * g1)) || (rc = grow(H->sv, sizeof(double4) * g1)) ||
        (rc = grow(H->terms, sizeof(HyTerms) * g1)) || (step && (rc = grow(H->sstep, 8 * g1))))
        return rc;
    hipLaunchKernelGGL(k_hy_keys, dim3(blocks_of(ngas, 256)), dim3(256), 0, st, gas, meta, x, y, z, minX, minY, minZ,
                       scale, (uint64_t*)H->keys.p, (int32_t*)H->vals.p);
    Meta* sm = (Meta*)H->sortMeta.p;
    hipLaunchKernelGGL(k_hy_sort_meta, dim3(1), dim3(1), 0, st, (const HydroMeta*)meta, sm);
    // stable LSD radix sort, 8 passes of 8 bits; ends in keys / vals
    Batch batch = {};
    batch.B = 1;
    uint64_t* kin = (uint64_t*)H->keys.p;
    uint64_t* kout = (uint64_t*)H->keys2.p;
    int32_t* vin = (int32_t*)H->vals.p;
    int32_t* vout = (int32_t*)H->vals2.p;
    for (int pass = 0; pass < 8; pass++) {
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(256), 0, st, kin, vin, sm, 8 * pass, 0, batch,
                           (uint32_t*)H->hist.p, ntiles);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kRowScanThreads), 0, st, (uint32_t*)H->hist.p, ntiles,
                           (uint32_t*)H->digitTotal.p);
        hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(256), 0, st, kin, vin, kout, vout, sm, 8 * pass, 0,
                           batch, (const uint32_t*)H->hist.p, (const uint32_t*)H->digitTotal.p, ntiles,
                           (int32_t*)nullptr);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    const int32_t* order = (const int32_t*)H->vals.p;
    double4* sp = (double4*)H->sp.p;
    hipLaunchKernelGGL(k_hy_gather, dim3(blocks_of(ngas, 256)), dim3(256), 0, st, order, meta, x, y, z,
                       (const double*)h, m, sp, (double*)H->sm.p);
    mark(1);
    // tree levels; the first batch covers the last tree's depth + 4
    BvhNode* nodes = (BvhNode*)H->nodes.p;
    QuadNode* quad = (QuadNode*)H->quad.p;
    int2* leafRange = (int2*)H->leafRange.p;
    hipLaunchKernelGGL(k_hy_root, dim3(1), dim3(1), 0, st, meta, nodes);
    int built = 0;
    int want = std::min(kBvhLevels, std::max(H->depth + 4, 24));
    while (true) {
        for (int L = built; L < want; L++) {
            const long long upper = std::min<long long>(L < 40 ? (1ll << L) : ngas, ngas);
            hipLaunchKernelGGL(k_hy_level, dim3((unsigned)std::min<long long>(1024, blocks_of(upper, 256))),
                               dim3(256), 0, st, L, (const uint64_t*)H->keys.p, nodes, meta, leafRange);
        }
        built = want;
        HYCHK(hipMemcpyAsync(H->hmeta, meta, sizeof(HydroMeta), hipMemcpyDeviceToHost, st));
        HYCHK(hipStreamSynchronize(st));
        if (H->hmeta->lvlCount[built] == 0) break;
        if (built == kBvhLevels)
            return set_error(GX_ERR_CAPACITY, "neighbour tree deeper than %d levels", kBvhLevels);
        want = std::min(kBvhLevels, built + 16);
    }
    int depth = 0;
    long long nn = 0;
    for (int L = 0; L < built; L++) {
        if (H->hmeta->lvlCount[L] == 0) break;
        depth = L + 1;
        nn += H->hmeta->lvlCount[L];
    }
    H->depth = depth;
    H->nnodes = nn;
    // Bottom-up refit: runs of levels with at most kSmallLevel nodes go to one
    // workgroup each, wider levels get a grid launch of their own.
    constexpr int kSmallLevel = 512;
    auto refit = [&](const double* hNew) {
        int L = depth - 1;
        while (L >= 0) {
            const long long cntL = H->hmeta->lvlCount[L];
            if (cntL > kSmallLevel) {
                hipLaunchKernelGGL(k_hy_refit, dim3((unsigned)std::min<long long>(1024, blocks_of(cntL, 256))),
                                   dim3(256), 0, st, L, (const double4*)sp, order, hNew, nodes,
                                   (const HydroMeta*)meta);
                L--;
                continue;
            }
            int lo = L;
            while (lo - 1 >= 0 && H->hmeta->lvlCount[lo - 1] <= kSmallLevel) lo--;
            hipLaunchKernelGGL(k_hy_refit_small, dim3(1), dim3(1024), 0, st, L, lo, (const double4*)sp, order, hNew,
                               nodes, (const HydroMeta*)meta);
            L = lo - 1;
        }
        hipLaunchKernelGGL(k_hy_quad, dim3(blocks_of(nn, 256)), dim3(256), 0, st, (const BvhNode*)nodes, nn, quad,
                           hNew ? (const int*)&meta->hChanged : nullptr);
    };
    // The walk stack holds at most three siblings per level of the 4-wide walk
    // plus the top.
    const int stackCap = 3 * ((depth + 1) / 2) + 4;
    const size_t walkLds = walk_lds_bytes(stackCap);
    if (walkLds > 64 * 1024) return set_error(GX_ERR_CAPACITY, "neighbour tree too deep (%d levels)", depth);
    refit(nullptr);
    mark(2);
    const long long nw = (ngas + kWave - 1) / kWave;
    const unsigned waveGrid = (unsigned)(8 * ((nw + 7) / 8));
    hipLaunchKernelGGL(k_hy_density, dim3(waveGrid), dim3(kWave), walkLds, st, (const QuadNode*)quad, stackCap,
                       (const double4*)sp,
                       (const double*)H->sm.p, order, active, meta, 0.25 * softening, ngb, h, rho);
    mark(3);
    refit(h);  // hmax of the final smoothing lengths (refitH; the kernels return at once if no h moved)
    hipLaunchKernelGGL(k_hy_sort_leaves, dim3(blocks_of(ngas, 256)), dim3(256), 0, st, (const HydroMeta*)meta,
                       (const int2*)leafRange, order, (const double4*)sp, (const double*)h, (int32_t*)H->order2.p,
                       (double4*)H->sp2.p);
    const int32_t* order2 = (const int32_t*)H->order2.p;
    const double4* sp2 = (const double4*)H->sp2.p;
    hipLaunchKernelGGL(k_hy_terms, dim3(blocks_of(ngas, 256)), dim3(256), 0, st, order2, (const HydroMeta*)meta, sp2,
                       vx, vy, vz, m, (const double*)rho, u, step, gamma, (double4*)H->sv.p, (HyTerms*)H->terms.p,
                       step ? (double*)H->sstep.p : nullptr);
    mark(4);
    hipLaunchKernelGGL(k_hy_force, dim3(waveGrid), dim3(kWave), walkLds, st, (const QuadNode*)quad, stackCap, sp2,
                       (const double4*)H->sv.p, (const HyTerms*)H->terms.p,
                       step ? (const double*)H->sstep.p : nullptr, order2, active, meta, alpha, ax, ay, az, dudt,
                       divv, dtOut, (unsigned long long*)wake);
    mark(5);
    HYCHK(hipGetLastError());
    HYCHK(hipMemcpyAsync(H->hmeta, meta, sizeof(HydroMeta), hipMemcpyDeviceToHost, st));
    HYCHK(hipStreamSynchronize(st));
    if (H->hmeta->flags) return set_error(GX_ERR_CAPACITY, "walk stack overflow (flags %d)", H->hmeta->flags);
    H->stats.n_nodes = nn;
    H->stats.n_leaves = H->hmeta->leaves;
    H->stats.depth = depth;
    H->stats.max_iterations = H->hmeta->maxPasses;
    H->stats.density_sums = (int64_t)H->hmeta->evals;
    H->stats.pairs = (int64_t)H->hmeta->pairs;
    H->stats.walk_steps = (int64_t)H->hmeta->walkSteps;
    H->stats.density_pairs = (int64_t)H->hmeta->densityPairs;
    H->stats.visc_pairs = (int64_t)H->hmeta->viscPairs;
    if (H->timing) {
        for (int p = 0; p < kHyPhases; p++) (void)hipEventElapsedTime(&H->phase[p], H->ev[p], H->ev[p + 1]);
        (void)hipEventElapsedTime(&H->phase[kHyPhases], H->ev[0], H->ev[kHyPhases]);
    }
    return GX_OK;
}

}  // namespace

extern "C" {

gx_hydro* gx_hydro_new(void) {
    if (device_ready() != GX_OK) return nullptr;
    gx_hydro* H = new gx_hydro();
    if (hipStreamCreateWithFlags(&H->own, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&H->hmeta, sizeof(HydroMeta), hipHostMallocDefault) != hipSuccess) {
        set_error(GX_ERR_HIP, "hydro stream or pinned allocation failed");
        gx_hydro_free(H);
        return nullptr;
    }
    memset(H->hmeta, 0, sizeof(HydroMeta));
    H->stream = H->own;
    for (auto& e : H->ev) (void)hipEventCreate(&e);
    return H;
}

void gx_hydro_free(gx_hydro* H) {
    if (!H) return;
    if (H->stream) (void)hipStreamSynchronize(H->stream);
    Buf* bufs[] = {&H->meta, &H->sortMeta, &H->flag, &H->tileCnt, &H->gas, &H->keys, &H->keys2, &H->vals,
                   &H->vals2, &H->hist, &H->digitTotal, &H->sp, &H->sp2, &H->sm, &H->order2, &H->nodes,
                   &H->leafRange, &H->sv, &H->terms, &H->sstep, &H->quad};
    for (Buf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (Buf& b : H->stage)
        if (b.p) (void)hipFree(b.p);
    for (auto& e : H->ev)
        if (e) (void)hipEventDestroy(e);
    if (H->hmeta) (void)hipHostFree(H->hmeta);
    if (H->own) (void)hipStreamDestroy(H->own);
    delete H;
}

int gx_hydro_config(gx_hydro* H, int timing) {
    if (!H) return set_error(GX_ERR_ARG, "null hydro");
    H->timing = timing != 0;
    return GX_OK;
}

int gx_hydro_set_stream(gx_hydro* H, void* stream) {
    if (!H) return set_error(GX_ERR_ARG, "null hydro");
    H->stream = stream ? (hipStream_t)stream : H->own;
    return GX_OK;
}

int gx_hydro_forces_dev(gx_hydro* H, const double* x, const double* y, const double* z, const double* vx,
                        const double* vy, const double* vz, const double* mass, const uint8_t* kind,
                        const uint8_t* alive, const double* u, double* hsml, double* rho, int64_t n,
                        const gx_gas_params* gas, double softening, const uint8_t* active, const double* step,
                        double* ax, double* ay, double* az, double* dudt, double* divv, double* dt,
                        uint64_t* wake) {
    if (!H) return set_error(GX_ERR_ARG, "null hydro");
    if (n < 0) return set_error(GX_ERR_ARG, "negative n");
    if (n >= (1LL << 26))  // walk links hold (node + 1) << 4 in an int32
        return set_error(GX_ERR_ARG, "n = %lld exceeds the hydro's limit of 2^26 bodies", (long long)n);
    if (n == 0) return GX_OK;
    if (!x || !y || !z || !vx || !vy || !vz || !mass || !kind || !alive || !u || !hsml || !rho || !ax || !ay ||
        !az || !dudt || !divv || !dt || !wake)
        return set_error(GX_ERR_ARG, "null array");
    return hydro_run(H, x, y, z, vx, vy, vz, mass, kind, alive, u, hsml, rho, n, gas, softening, active, step, ax, ay,
                     az, dudt, divv, dt, wake);
}

int gx_hydro_forces(gx_hydro* H, const double* x, const double* y, const double* z, const double* vx,
                    const double* vy, const double* vz, const double* mass, const uint8_t* kind,
                    const uint8_t* alive, const double* u, double* hsml, double* rho, int64_t n,
                    const gx_gas_params* gas, double softening, const uint8_t* active, const double* step,
                    double* ax, double* ay, double* az, double* dudt, double* divv, double* dt, uint64_t* wake) {
    if (!H) return set_error(GX_ERR_ARG, "null hydro");
    if (n < 0) return set_error(GX_ERR_ARG, "negative n");
    if (n >= (1LL << 26))
        return set_error(GX_ERR_ARG, "n = %lld exceeds the hydro's limit of 2^26 bodies", (long long)n);
    if (n == 0) return GX_OK;
    if (!x || !y || !z || !vx || !vy || !vz || !mass || !kind || !alive || !u || !hsml || !rho || !ax || !ay ||
        !az || !dudt || !divv || !dt || !wake)
        return set_error(GX_ERR_ARG, "null array");
    const size_t d = 8 * (size_t)n, b = (size_t)n;
    // staging: 0..6 x y z vx vy vz m, 7 kind, 8 alive, 9 u, 10 h, 11 rho, 12 active, 13 step,
    // 14..19 ax ay az dudt divv dt, 20 wake
    const void* in[] = {x, y, z, vx, vy, vz, mass};
    int rc;
    for (int s = 0; s < 21; s++) {
        const size_t bytes = (s == 7 || s == 8 || s == 12) ? b : d;
        if ((rc = grow(H->stage[s], bytes))) return rc;
    }
    Buf& wakeBuf = H->stage[20];
    hipStream_t st = H->stream;
    for (int s = 0; s < 7; s++) HYCHK(hipMemcpyAsync(H->stage[s].p, in[s], d, hipMemcpyHostToDevice, st));
    HYCHK(hipMemcpyAsync(H->stage[7].p, kind, b, hipMemcpyHostToDevice, st));
    HYCHK(hipMemcpyAsync(H->stage[8].p, alive, b, hipMemcpyHostToDevice, st));
    HYCHK(hipMemcpyAsync(H->stage[9].p, u, d, hipMemcpyHostToDevice, st));
    HYCHK(hipMemcpyAsync(H->stage[10].p, hsml, d, hipMemcpyHostToDevice, st));
    HYCHK(hipMemcpyAsync(H->stage[11].p, rho, d, hipMemcpyHostToDevice, st));
    if (active) HYCHK(hipMemcpyAsync(H->stage[12].p, active, b, hipMemcpyHostToDevice, st));
    if (step) HYCHK(hipMemcpyAsync(H->stage[13].p, step, d, hipMemcpyHostToDevice, st));
    double* out[6] = {ax, ay, az, dudt, divv, dt};
    for (int s = 0; s < 6; s++) HYCHK(hipMemcpyAsync(H->stage[14 + s].p, out[s], d, hipMemcpyHostToDevice, st));
    HYCHK(hipMemcpyAsync(wakeBuf.p, wake, d, hipMemcpyHostToDevice, st));
    rc = hydro_run(H, (double*)H->stage[0].p, (double*)H->stage[1].p, (double*)H->stage[2].p, (double*)H->stage[3].p,
                   (double*)H->stage[4].p, (double*)H->stage[5].p, (double*)H->stage[6].p, (uint8_t*)H->stage[7].p,
                   (uint8_t*)H->stage[8].p, (double*)H->stage[9].p, (double*)H->stage[10].p,
                   (double*)H->stage[11].p, n, gas, softening, active ? (uint8_t*)H->stage[12].p : nullptr,
                   step ? (double*)H->stage[13].p : nullptr, (double*)H->stage[14].p, (double*)H->stage[15].p,
                   (double*)H->stage[16].p, (double*)H->stage[17].p, (double*)H->stage[18].p,
                   (double*)H->stage[19].p, (uint64_t*)wakeBuf.p);
    if (rc == GX_OK) {
        HYCHK(hipMemcpyAsync(hsml, H->stage[10].p, d, hipMemcpyDeviceToHost, st));
        HYCHK(hipMemcpyAsync(rho, H->stage[11].p, d, hipMemcpyDeviceToHost, st));
        for (int s = 0; s < 6; s++) HYCHK(hipMemcpyAsync(out[s], H->stage[14 + s].p, d, hipMemcpyDeviceToHost, st));
        HYCHK(hipMemcpyAsync(wake, wakeBuf.p, d, hipMemcpyDeviceToHost, st));
        HYCHK(hipStreamSynchronize(st));
    }
    (void)hipStreamSynchronize(st);
    return rc;
}

int gx_hydro_get_stats(const gx_hydro* H, gx_hydro_stats* out) {
    if (!H || !out) return set_error(GX_ERR_ARG, "null argument");
    *out = H->stats;
    return GX_OK;
}

int gx_hydro_phase_ms(const gx_hydro* H, float* ms, int n) {
    if (!H || !ms) return set_error(GX_ERR_ARG, "null argument");
    const int k = std::min(n, kHyPhases + 1);
    for (int p = 0; p < k; p++) ms[p] = H->timing ? H->phase[p] : 0.f;
    return k;
}

int64_t gx_hydro_export(gx_hydro* H, int what, void* host, int64_t count) {
    if (!H || !host || count < 0) return set_error(GX_ERR_ARG, "bad export arguments");
    const void* src;
    size_t item;
    long long avail;
    switch (what) {
        case 0: src = H->keys.p; item = 8; avail = H->ngas; break;
        case 1: src = H->order2.p; item = 4; avail = H->ngas; break;
        case 2: src = H->nodes.p; item = sizeof(BvhNode); avail = H->nnodes; break;
        default: return set_error(GX_ERR_ARG, "unknown export %d", what);
    }
    if (H->ngas < 2 || !src) return 0;
    const long long c = std::min<long long>(count, avail);
    if (c > 0) {
        if (hipMemcpy(host, src, item * (size_t)c, hipMemcpyDeviceToHost) != hipSuccess)
            return set_error(GX_ERR_HIP, "export copy failed");
    }
    return c;
}

}  // extern "C"
