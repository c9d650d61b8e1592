// gx_common.h -- device records and wavefront helpers shared by the gfx950
// Barnes-Hut kernels.  All kernels are written for CDNA4 wave64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gx {

constexpr int kMortonBits = 21;   // internal/sim/bh/morton.go:4
constexpr int kLeafSize = 24;     // internal/sim/bh/bh.go:19
constexpr int kMaxLevels = 22;    // levels 0..21
constexpr int kWave = 64;

enum Kind : uint8_t { kStar = 0, kHole = 1, kDark = 2, kGas = 3 };

// Per-node traversal record (64 B, one cache line).  During the topology
// build x,y,z hold the cell centre and crit2 holds the half side; the moment
// pass replaces them with the centre of mass and the squared opening radius.
struct alignas(64) TravNode {
    double x, y, z, crit2;
    int32_t first;   // first child, -1 for leaves
    int32_t next;    // skip link: next sibling or the parent's next, -1 at the end
    int32_t lo, hi;  // star range in sorted order
    int32_t nchild;  // children first .. first+nchild-1 (octant order)
    int32_t level;
    int32_t parent;  // -1 for a root
    int32_t inst;    // instance (system) the node belongs to
};

// FP64 multipole (m and the traceless quadrupole about the centre of mass).
struct alignas(64) Mult64 {
    double m, qxx, qyy, qzz, qxy, qxz, qyz, pad;
};

// FP32 multipole for the FP32 walk.
struct alignas(32) Mult32 {
    float m, qxx, qyy, qzz, qxy, qxz, qyz, pad;
};

// FP16-stored multipole: m in FP32, the quadrupole scaled by 1/(m h^2) with h
// the cell's half side (|Q_ij| <= 36 m h^2 because every star of a cell lies
// within 2 sqrt(3) h of its centre of mass) and rounded to FP16.
struct alignas(16) Mult16 {
    float m;
    uint16_t q[6];  // qxx, qyy, qzz, qxy, qxz, qyz as IEEE half
};

// Instances of one step: B independent systems packed one after the other in
// the body arrays (a single system is B = 1).  Each gets its own bounding box,
// root cell, tree, opening angle and softening; walk groups never mix them.
struct Batch {
    int B;
    int gsize;                      // stars per walk group (32 or 64)
    const long long* bodyOff;       // B+1 body offsets in the arena
    const double* theta;            // B opening angles (DefaultTheta already applied)
    const double* eps2;             // B squared softenings
    unsigned long long* boxEnc;     // 6B: encoded min x,y,z, max x,y,z of the instance's stars
    unsigned long long* starCount;  // B
    long long* starOff;             // B+1 offsets of the instances in sorted order
    long long* groupOff;            // B+1 offsets of the instances' walk groups
    int* root;                      // B root node, -1 for an instance without stars
    double* half;                   // B half side of the root cell
    // Targets of the walk: positions [tOff[b], tOff[b+1]) of the target list
    // belong to instance b, and target t is the sorted star tlist[t].  Without
    // an active set tlist is null, tOff is starOff and target t is star t.
    const long long* tOff;
    const int32_t* tlist;
    // Instance of each body (B > 1; written by the collect for the sort's
    // instance passes), null for a single system.
    const uint16_t* instOf;
};

// Sorted star index of walk target t.
__device__ __forceinline__ long long target_star(const Batch& b, long long t) { return b.tlist ? b.tlist[t] : t; }

// Order-preserving map of doubles to unsigned integers, so that atomic
// integer min/max give exact floating-point min/max (-0 orders below +0).
__device__ __forceinline__ unsigned long long enc_d(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dec_d(unsigned long long u) {
    return __longlong_as_double((long long)((u >> 63) ? (u & 0x7fffffffffffffffull) : ~u));
}

// Largest i with off[i] <= v among off[0..count) (off ascending, off[0] <= v).
__device__ __forceinline__ int upper_index(const long long* __restrict__ off, int count, long long v) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Device-side metadata of one build.
struct Meta {
    long long nstars;
    long long nholes;
    long long nnodes;
    long long levelStart[kMaxLevels + 1];
    long long levelCount[kMaxLevels + 1];
    unsigned long long pp;  // particle-particle interactions (incl. self pairs)
    unsigned long long pc;  // particle-cell interactions
    unsigned long long xcdNext[8];   // work queues of the persistent walk, one per XCD
    unsigned long long listEntries;  // cell list entries over all groups (FP32 walk)
    unsigned long long rounds;       // FP32 walk traversal rounds over all groups
    unsigned long long mixed;        // nodes decided per star (box test inconclusive)
    long long ntargets;     // walk targets (the active stars of an active set)
    int flags;              // Flag bits
    int maxStack;
    int maxList[3];         // longest full-cell, partial-cell and leaf list of a group
    // Flags of every step since the host last read them: each force step ORs
    // its flags in at its end (k_zero_holes), k_step_init leaves it alone, and
    // the host clears it once it has reported it.  Replays of a captured step
    // thus keep an overflow of an earlier replay.
    unsigned long long sticky;
    // kFlagTopology: which link checks failed (OR of 1 record index, 2 record
    // fields, 4 end, 8 skip link, 16 first child, 32 split at the last level,
    // 64 keys out of order)
    // and the lowest node that failed one
    int topoWhy;
    long long topoNode;
};

enum Flag : int {
    kFlagNodeOverflow = 1,   // node arena too small
    kFlagStackOverflow = 2,  // FP32 walk stack too small
    kFlagListOverflow = 4,   // FP32 walk lists too small
    kFlagDepth = 8,          // the tree needs more levels than were built
    kFlagTopology = 16,      // a node's links failed their check (never expected)
};
// The tree of this step is incomplete: later kernels skip their work.
constexpr int kFlagTreeIncomplete = kFlagNodeOverflow | kFlagDepth | kFlagTopology;

// ---------------------------------------------------------------------------
// wavefront helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, kWave); }

// A double that is the same in every active lane, moved to scalar registers.
__device__ __forceinline__ double uniform_d(double v) {
    const long long u = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __longlong_as_double((long long)((uint64_t)hi << 32 | lo));
}

__device__ __forceinline__ double wave_min_d(double v) {
    for (int m = 32; m >= 1; m >>= 1) v = fmin(v, shfl_xor_d(v, m));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, shfl_xor_d(v, m));
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}
// Inclusive prefix sum across the wave (every lane active), with DPP moves:
// shifts by 1, 2, 4 and 8 within each row of 16 lanes, then the last lane of
// rows 0 and 2 added to rows 1 and 3, and the last lane of row 1 to rows 2
// and 3.  Lanes a move has no source for add the old value 0.
__device__ __forceinline__ int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return v;
}

// v_writelane_b32: lane `lane` of the result is `value` (uniform), the other
// lanes keep `old` (the LLVM intrinsic, which has no clang builtin here).
extern "C" __device__ int gx_llvm_writelane(int value, int lane, int old) __asm("llvm.amdgcn.writelane");

// Inclusive running maximum across the wave (every lane active, values >= 0),
// with the same DPP moves as wave_incl_scan.
__device__ __forceinline__ int wave_incl_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15 into rows 1, 3
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31 into rows 2, 3
    return v;
}

// Morton interleave: morton.go:7-21.
__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= (1ull << kMortonBits) - 1;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

}  // namespace gx
