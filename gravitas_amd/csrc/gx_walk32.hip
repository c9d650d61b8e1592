// gx_walk32.hip -- FP32 group walk for gfx950 (wave64).
//
// One wavefront walks the tree for a group of 64 consecutive targets in
// Morton order (one star per lane), in two phases per group:
//
// Traversal (trav_group): a shared stack of child ranges in LDS, each entry
// with a 64-bit lane mask.  Up to 64 nodes are popped per round and tested
// against the group's bounding box in FP64; when the box does not decide a
// node (some stars would accept it and others open it), the reference's own
// per-star test d2 >= crit2 (bh.go:114-116) is evaluated in FP64 by every
// lane and collected with a ballot.  The interaction set of every star is
// therefore exactly the one of the reference's per-particle walk
// (bh.go:106-155); only the arithmetic of the interactions is FP32.  The
// group's lists go to HBM: cells accepted by all lanes, cells accepted by
// some lanes and opened leaves (with their lane masks).
//
// Evaluation (eval_group): cells are staged 64 at a time in LDS as
// (COM - c_g, m, Q) in FP32, stars of opened leaves PARTS at a time as
// (x - c_g, m), where c_g is the group's bounding-box centre; coordinates
// are made group-relative in FP64 before rounding, so FP32 only carries
// offsets that are small for near interactions.  Entries that apply to every
// lane are evaluated by all lanes in a uniform loop; for the others the 64
// lane masks of a chunk are transposed, and each lane loops over its own
// entries only.  Partial-mask entries are dealt to chunks (cells) and flushes
// (leaves) with a stride across the whole list: neighbouring entries have
// similar masks, so dealing them apart evens out the per-lane loop lengths
// of each chunk or flush.
//
// Interactions follow bh.go:117-133 (cells, monopole + quadrupole) and
// bh.go:145-150 (stars, Plummer softening).  Each chunk is summed in FP32 and
// added to the per-star total with Kahan compensation.
#include <hip/hip_fp16.h>

#include "gx_common.h"
#include "gx_kernels.h"

namespace gx {

constexpr int kChunk = 64;    // cells staged per evaluation chunk
constexpr double kMargin = 1e-12;  // >> FP64 rounding of d2; box tests stay conservative

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// One cell's pull on a star at (px, py, pz), group-relative: monopole and
// quadrupole terms of bh.go:117-133 with (c0.xyz, c0.w) = (COM, m) and
// (c1, c2.xy) = (qxx, qyy, qzz, qxy, qxz, qyz).
template <bool QUAD>
__device__ __forceinline__ void cell_force(const float4 c0, const float4 c1, const float4 c2, float px, float py,
                                           float pz, float eps2, float& ax, float& ay, float& az) {
    const float dx = c0.x - px, dy = c0.y - py, dz = c0.z - pz;
    const float d2e = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    const float ir = __builtin_amdgcn_rsqf(d2e);
    const float inv = ir * ir;
    const float inv3 = inv * ir;
    if (QUAD) {
        // With e = d inv and g = Q e = inv Qd:  f d - Qd inv5
        //   = inv3 ((m + 2.5 (e.g)) d - g),  since e.g = inv^2 d.Qd;
        // 31 instructions against 32 for the term-by-term form.
        const float ex = dx * inv, ey = dy * inv, ez = dz * inv;
        const float gx = fmaf(c1.x, ex, fmaf(c1.w, ey, c2.x * ez));
        const float gy = fmaf(c1.w, ex, fmaf(c1.y, ey, c2.y * ez));
        const float gz = fmaf(c2.x, ex, fmaf(c2.y, ey, c1.z * ez));
        const float s = fmaf(2.5f, fmaf(ex, gx, fmaf(ey, gy, ez * gz)), c0.w);
        ax = fmaf(inv3, fmaf(s, dx, -gx), ax);
        ay = fmaf(inv3, fmaf(s, dy, -gy), ay);
        az = fmaf(inv3, fmaf(s, dz, -gz), az);
    } else {
        const float f = c0.w * inv3;
        ax = fmaf(f, dx, ax);
        ay = fmaf(f, dy, ay);
        az = fmaf(f, dz, az);
    }
}

// Cells accepted by every lane of the group: all lanes evaluate every cell.
template <bool QUAD>
__device__ __forceinline__ void eval_cells(const float4* __restrict__ cell, int ncell, float px, float py, float pz,
                                           float eps2, float& ax, float& ay, float& az) {
    ncell = __builtin_amdgcn_readfirstlane(ncell);  // uniform: the loop counter lives in SGPRs
#pragma unroll 4
    for (int e = 0; e < ncell; e++)
        cell_force<QUAD>(cell[3 * e], cell[3 * e + 1], cell[3 * e + 2], px, py, pz, eps2, ax, ay, az);
}

// One butterfly stage of the bit-matrix transposes on a 32-bit word: lanes l
// and l ^ s swap blocks of s bits.  The lower lane keeps the blocks of m and
// sends its other blocks down by s; the upper lane keeps the other blocks
// and sends the blocks of m up by s.  A shift of the sent blocks never
// carries a bit past the word's end, so it is a rotation by a per-lane
// amount (one v_alignbit_b32), and the lane's own part is w ^ t.
struct Stage {
    uint32_t nm, rot;  // blocks to send, rotation right that moves them
    __device__ __forceinline__ Stage(int lane, int s, uint32_t m)
        : nm((lane & s) ? m : ~m), rot((lane & s) ? 32 - s : s) {}
    __device__ __forceinline__ uint32_t apply(uint32_t w, int s) const {
        const uint32_t t = w & nm;
        const uint32_t send = __builtin_amdgcn_alignbit(t, t, rot);
        return (w ^ t) | (uint32_t)__shfl_xor((int)send, s, kWave);
    }
};

// Transpose of the 64 x 64 bit matrix whose row l is held by lane l: returns
// column l (bit e set iff bit l of lane e's row is set).  Six butterfly
// stages: the first exchanges whole words (the lower half of the lanes keeps
// its low words and takes its partners' low words as high words, the upper
// half the reverse), the other five run on both words with the same blocks.
__device__ __forceinline__ uint64_t transpose_bits(uint64_t x) {
    const int lane = lane_id();
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    {
        const bool up = (lane & 32) != 0;
        const uint32_t recv = (uint32_t)__shfl_xor((int)(up ? lo : hi), 32, kWave);
        if (up)
            lo = recv;
        else
            hi = recv;
    }
    constexpr uint32_t keep[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int sft = 16 >> k;
        const Stage st(lane, sft, keep[k]);
        lo = st.apply(lo, sft);
        hi = st.apply(hi, sft);
    }
    return (uint64_t)hi << 32 | lo;
}

// The same for a 32 x 64 matrix whose rows sit on lanes 0-31 (lanes 32-63
// pass 0): column l has 32 bits.  The first stage only moves the high words
// of lanes 0-31 to lanes 32-63; the other five run on 32-bit words.
__device__ __forceinline__ uint32_t transpose_bits32(uint64_t x) {
    const int lane = lane_id();
    const uint32_t hi = __shfl_xor((int)(uint32_t)(x >> 32), 32, kWave);
    uint32_t w = lane < 32 ? (uint32_t)x : hi;
    constexpr uint32_t keep[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int sft = 16 >> k;
        w = Stage(lane, sft, keep[k]).apply(w, sft);
    }
    return w;
}

// Cells accepted by some lanes only (mask of cell e held by lane e): each
// lane loops over its own cells, so no lane slot is spent on a masked-off
// star.  col = transpose of the masks.
// The column is walked as two 32-bit words (cells 0-31, then 32-63): a 32-bit
// find-first-bit, clear and test cost four VALU instructions per cell against
// nine for the 64-bit word.
template <bool QUAD>
__device__ __forceinline__ void eval_cells_own(const float4* __restrict__ cell, uint64_t col, float px, float py,
                                               float pz, float eps2, float& ax, float& ay, float& az) {
#pragma unroll
    for (int half = 0; half < 2; half++) {
        uint32_t w = (uint32_t)(col >> (32 * half));
        const float4* __restrict__ c = cell + 96 * half;
        while (w) {
            const int e = __builtin_ctz(w);
            w &= w - 1;
            cell_force<QUAD>(c[3 * e], c[3 * e + 1], c[3 * e + 2], px, py, pz, eps2, ax, ay, az);
        }
    }
}

// One star's pull (bh.go:145-150, Plummer softening); with EPS0 the star
// itself (d = 0 exactly) is skipped.
template <bool EPS0>
__device__ __forceinline__ void star_force(const float4 q, float px, float py, float pz, float eps2, float& ax,
                                           float& ay, float& az) {
    const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    float ir = __builtin_amdgcn_rsqf(r2);
    if (EPS0) ir = r2 > 0.0f ? ir : 0.0f;
    const float mir = q.w * ir;
    const float f = mir * (ir * ir);
    ax = fmaf(f, dx, ax);
    ay = fmaf(f, dy, ay);
    az = fmaf(f, dz, az);
}

// Stars [first, first + npart) of the staging buffer, for every lane.
template <bool EPS0>
__device__ __forceinline__ void eval_parts(const float4* __restrict__ part, int first, int npart, float px, float py,
                                           float pz, float eps2, float& ax, float& ay, float& az) {
    first = __builtin_amdgcn_readfirstlane(first);
    const int end = first + __builtin_amdgcn_readfirstlane(npart);
#pragma unroll 2
    for (int e = first; e < end; e++) star_force<EPS0>(part[e], px, py, pz, eps2, ax, ay, az);
}

// Stars [first, first + npart) of the staging buffer, each for the lanes of
// its leaf's mask rmask[range[]]: blocks of 32 stars, masks transposed (lanes 0-31
// hold the rows, so every column fits in 32 bits), and each lane loops over
// its own stars.  A 32-bit column costs four VALU instructions of loop
// control per star against nine for a 64-bit one; the lane that does the most
// work is mostly the same in every block, so smaller blocks add little
// imbalance.
template <bool EPS0>
__device__ __forceinline__ void eval_parts_own(const float4* __restrict__ part, const uint8_t* __restrict__ range,
                                               const uint64_t* __restrict__ rmask, int first, int npart, float px,
                                               float py, float pz, float eps2, float& ax, float& ay, float& az) {
    constexpr int kBlk = 32;
    first = __builtin_amdgcn_readfirstlane(first);
    npart = __builtin_amdgcn_readfirstlane(npart);
    // block b holds stars b, b + nb, b + 2 nb, ... (spread over the leaves)
    const int nb = (npart + kBlk - 1) / kBlk;
    const int lane = lane_id();
    for (int b = 0; b < nb; b++) {
        const int mine = b + nb * lane;
        uint32_t col = transpose_bits32(lane < kBlk && mine < npart ? rmask[range[first + mine]] : 0ull);
        const float4* blk = part + first + b;
        while (col) {
            const int e = __builtin_ctz(col);
            col &= col - 1;
            star_force<EPS0>(blk[__umul24(nb, e)], px, py, pz, eps2, ax, ay, az);
        }
    }
}

// Element i (0 <= i < 2^24, the list capacities' bound) of an array whose base
// is uniform, through a 32-bit byte offset: a scalar base plus one vector
// offset register from a 24-bit multiply, no 64-bit address math.
template <typename T>
__device__ __forceinline__ T& at32(T* base, int i) {
    return *(T*)((char*)base + __umul24((uint32_t)i, (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T& at32(const T* base, int i) {
    return *(const T*)((const char*)base + __umul24((uint32_t)i, (uint32_t)sizeof(T)));
}

// Group prologue shared by both kernels: the group's stars (up to gsize
// consecutive targets of one instance in Morton order: all its stars, or the
// active ones of an active set), FP64 bounding box and centre c_g, the
// instance's root, softening and root half side.
struct Group {
    long long g, k;
    int inst, root;
    bool valid;
    float eps2;
    double half;
    double4 p;
    uint64_t gmask;
    double bminx, bminy, bminz, bmaxx, bmaxy, bmaxz, cgx, cgy, cgz;
};

__device__ __forceinline__ bool load_group(Group& G, const double4* __restrict__ spos, const Batch& batch,
                                           long long g) {
    if (g >= batch.groupOff[batch.B]) return false;
    G.g = g;
    G.inst = upper_index(batch.groupOff, batch.B, g);
    const long long base = batch.tOff[G.inst] + (g - batch.groupOff[G.inst]) * batch.gsize;
    const long long count = min((long long)batch.gsize, batch.tOff[G.inst + 1] - base);
    G.root = __builtin_amdgcn_readfirstlane(batch.root[G.inst]);
    G.eps2 = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)batch.eps2[G.inst])));
    G.half = uniform_d(batch.half[G.inst]);
    G.valid = lane_id() < count;
    G.k = target_star(batch, G.valid ? base + lane_id() : base);
    G.p = spos[G.k];
    G.gmask = __ballot(G.valid);
    // the box and its centre are the same in every lane: scalar registers
    G.bminx = uniform_d(wave_min_d(G.valid ? G.p.x : INFINITY));
    G.bmaxx = uniform_d(wave_max_d(G.valid ? G.p.x : -INFINITY));
    G.bminy = uniform_d(wave_min_d(G.valid ? G.p.y : INFINITY));
    G.bmaxy = uniform_d(wave_max_d(G.valid ? G.p.y : -INFINITY));
    G.bminz = uniform_d(wave_min_d(G.valid ? G.p.z : INFINITY));
    G.bmaxz = uniform_d(wave_max_d(G.valid ? G.p.z : -INFINITY));
    G.cgx = uniform_d(0.5 * (G.bminx + G.bmaxx));
    G.cgy = uniform_d(0.5 * (G.bminy + G.bmaxy));
    G.cgz = uniform_d(0.5 * (G.bminz + G.bmaxz));
    return true;
}

// Group of this workgroup: workgroups are dealt round-robin to the 8 XCDs, so
// each XCD gets a contiguous run of (spatially adjacent) groups for its L2.
// The grid is an upper bound; the run length comes from the actual group
// count, so a small active set still spreads over all XCDs.  Workgroups past
// 8 runs get a group index past the end and exit.
__device__ __forceinline__ long long xcd_group(const Batch& batch) {
    const long long ng = batch.groupOff[batch.B];
    const long long per = (ng + 7) / 8;
    const long long j = blockIdx.x / 8;
    return j < per ? (long long)(blockIdx.x % 8) * per + j : ng;
}

// ---------------------------------------------------------------------------
// traversal: one wavefront per group, lists to HBM
// ---------------------------------------------------------------------------
template <int STACK>
struct TravLDS {
    uint64_t stkMask[STACK];
    uint32_t stkRange[STACK];  // first | (count - 1) << 28
    int popOwner[kWave];       // lane whose range the popped slot starts
};

// The stars of opened leaves staged for a flush, the staging tables of up to
// 64 leaf ranges (lane mask, exclusive prefix of the counts and first star of
// each range), and for the stars of partly opened leaves the range each came
// from (its lane mask is the range's: a byte per star instead of the mask).
template <int PARTS>
struct StarStage {
    float4 part[PARTS];
    uint64_t tMask[kWave];
    int tStart[kWave];
    int tCur[kWave];
    uint8_t range[PARTS];
};

// A group evaluates its cells first, then its stars: the cell chunk and the
// star staging share the memory.
template <int PARTS>
union EvalLDS {
    float4 cell[kChunk * 3];
    StarStage<PARTS> s;
};

template <int STACK, int PARTS>
union WalkLDS {
    TravLDS<STACK> t;
    EvalLDS<PARTS> e;
};

// Writes group G's interaction lists into region r; returns false on a stack
// or list overflow.
template <int STACK>
__device__ bool trav_group(const Group& G, long long r, TravLDS<STACK>& S, const TravNode* __restrict__ trav,
                           Meta* meta, const Lists& lists) {
    uint64_t* stkMask = S.stkMask;
    uint32_t* stkRange = S.stkRange;
    int* popOwner = S.popOwner;
    const int lane = lane_id();
    const double4 p = G.p;
    uint32_t* __restrict__ fullOut = lists.full + (size_t)r * lists.cf;
    uint3* __restrict__ partOut = lists.part + (size_t)r * lists.cp;
    uint3* __restrict__ leafOut = lists.leaf + (size_t)r * lists.cl;
    unsigned long long npp = 0, npc = 0;
    int nf = 0, np = 0, nlf = 0, nlp = 0, nps = 0, sp = 1, maxsp = 1, nrounds = 0, nmixed = 0;
    bool overflow = false;
    if (lane == 0) {
        stkRange[0] = (uint32_t)G.root;  // the instance's root, one node
        stkMask[0] = G.gmask;
    }
    __syncthreads();

    while (sp > 0) {
        // pop up to 64 nodes from the top ranges (top entry first)
        int cnt = 0;
        uint32_t rng = 0;
        uint64_t rmask = 0;
        if (lane < sp) {
            rng = stkRange[sp - 1 - lane];
            rmask = stkMask[sp - 1 - lane];
            cnt = (int)(rng >> 28) + 1;
        }
        const int incl0 = wave_incl_scan(cnt);
        const int excl0 = incl0 - cnt;
        const int take = max(0, min(cnt, kWave - excl0));
        const int rfirst = (int)(rng & 0x0fffffffu);
        const int nfull = __popcll(__ballot(cnt > 0 && take == cnt));
        const int npop = min(kWave, __shfl(incl0, 63, kWave));
        // a partly taken range stays on the stack with its remaining children
        if (lane == nfull && take > 0 && take < cnt)
            stkRange[sp - 1 - lane] = (uint32_t)(rfirst + take) | ((uint32_t)(cnt - take - 1) << 28);
        sp -= nfull;
        nrounds++;
        // Slot j of the round holds node rfirst + (j - excl0) of the last lane
        // whose run starts at or before j: each taking lane marks its first
        // slot, and a running maximum over the slots spreads the marks (the
        // lanes' runs ascend with the lane), so no lane loops over its run.
        popOwner[lane] = 0;
        __syncthreads();
        if (take > 0) popOwner[excl0] = lane;
        __syncthreads();
        const int owner = wave_incl_max(popOwner[lane]);
        const bool has = lane < npop;
        const int ofirst = __shfl(rfirst, owner, kWave);
        const int ostart = __shfl(excl0, owner, kWave);
        const uint64_t omaskAll = (uint64_t)__shfl((long long)rmask, owner, kWave);
        int node = 0;
        uint64_t mask = 0;
        if (has) {
            node = ofirst + (lane - ostart);
            mask = omaskAll;
        }
        int tfirst = -1, tnchild = 0, tlo = 0, thi = 0;
        int cls = 0;  // 1 accept for all, 2 open for all, 3 decide per star
        if (has) {
            const TravNode t = trav[node];
            tfirst = t.first; tnchild = t.nchild; tlo = t.lo; thi = t.hi;
            // nearest and farthest distance from the group's box to the
            // node's centre of mass; this classification is the walk's own
            // (the margin keeps it conservative, fused or not), only the
            // per-star test below must be the reference's arithmetic
            const double lx = G.bminx - t.x, hx = t.x - G.bmaxx;
            const double ly = G.bminy - t.y, hy = t.y - G.bmaxy;
            const double lz = G.bminz - t.z, hz = t.z - G.bmaxz;
            const double ex = fmax(0.0, fmax(lx, hx)), ey = fmax(0.0, fmax(ly, hy)), ez = fmax(0.0, fmax(lz, hz));
            const double fx = fmax(fabs(lx), fabs(hx)), fy = fmax(fabs(ly), fabs(hy)), fz = fmax(fabs(lz), fabs(hz));
            const double dmin2 = fma(ex, ex, fma(ey, ey, ez * ez));
            const double dmax2 = fma(fx, fx, fma(fy, fy, fz * fz));
            cls = dmin2 >= t.crit2 * (1 + kMargin) ? 1 : dmax2 * (1 + kMargin) < t.crit2 ? 2 : 3;
        }
        const uint64_t mixed = __ballot(cls == 3);
        nmixed += __popcll(mixed);
        // lane src: the ballot of its mixed node, written into that lane from
        // the scalar registers (two writelanes per node)
        // (a node every star accepts starts from an all-ones ballot, one they
        // all open from zero)
        int accLo = cls == 1 ? -1 : 0, accHi = accLo;
        for (uint64_t todo = mixed; todo;) {
            const int src = __builtin_ctzll(todo);
            todo &= todo - 1;
            // the node is uniform across the wave: its record comes in through a
            // scalar load rather than eight lane reads
            const int nid = __builtin_amdgcn_readlane(node, src);
            const double nx = trav[nid].x, ny = trav[nid].y, nz = trav[nid].z;
            const double ncrit2 = trav[nid].crit2;
            // the reference's per-star opening test, bh.go:114-116 (FP64, unfused)
            const double dx = nx - p.x, dy = ny - p.y, dz = nz - p.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const uint64_t acc = __ballot(d2 >= ncrit2);
            // lane src keeps the ballot and applies its node's lane mask after
            // the loop (no lane reads of the mask)
            accLo = gx_llvm_writelane((int)(uint32_t)acc, src, accLo);
            accHi = gx_llvm_writelane((int)(uint32_t)(acc >> 32), src, accHi);
        }
        const uint64_t acc = (uint64_t)(uint32_t)accHi << 32 | (uint32_t)accLo;
        const uint64_t amask = acc & mask, omask = mask & ~acc;

        // The classes of the round's nodes as wave masks (compares straight
        // into scalar registers; each lane's class is read back from the mask
        // by the branch itself, with no vector instruction).
        const uint64_t bcell = __ballot(amask != 0);
        const uint64_t bfull = __ballot(amask == G.gmask) & bcell;
        const uint64_t bpart = bcell & ~bfull;
        const uint64_t bopen = __ballot(omask != 0);
        const uint64_t bleaf = __ballot(tfirst < 0) & bopen;
        const uint64_t blf = __ballot(omask == G.gmask) & bleaf;
        const uint64_t blp = bleaf & ~blf;
        const uint64_t pb = bopen & ~bleaf;  // opened internal nodes

        // accepted cells -> full or partial list
        if (__builtin_amdgcn_inverse_ballot_w64(bfull)) {
            const int pos = nf + mbcnt(bfull);
            if (pos < lists.cf) at32(fullOut, pos) = (uint32_t)node;
        }
        if (__builtin_amdgcn_inverse_ballot_w64(bpart)) {
            const int pos = np + mbcnt(bpart);
            if (pos < lists.cp) at32(partOut, pos) = make_uint3((uint32_t)amask, (uint32_t)(amask >> 32), (uint32_t)node);
        }
        if (__builtin_amdgcn_inverse_ballot_w64(bcell)) npc += (unsigned long long)__popcll(amask);
        nf += __popcll(bfull);
        np += __popcll(bpart);

        // opened leaves -> leaf list: opened by every star from the front,
        // by some from the back
        if (__builtin_amdgcn_inverse_ballot_w64(blf)) {
            const int pos = nlf + mbcnt(blf);
            if (pos < lists.cl) at32(leafOut, pos) = make_uint3((uint32_t)omask, (uint32_t)(omask >> 32), (uint32_t)node);
        }
        if (__builtin_amdgcn_inverse_ballot_w64(blp)) {
            const int pos = nlp + mbcnt(blp);
            if (pos < lists.cl)
                at32(leafOut, lists.cl - 1 - pos) = make_uint3((uint32_t)omask, (uint32_t)(omask >> 32), (uint32_t)node);
            nps += thi - tlo;
        }
        if (__builtin_amdgcn_inverse_ballot_w64(bleaf))
            npp += (unsigned long long)(uint32_t)__popcll(omask) * (uint32_t)(thi - tlo);
        nlf += __popcll(blf);
        nlp += __popcll(blp);

        // opened internal nodes -> one stack entry with the whole child range
        const int npush = __popcll(pb);
        if (sp + npush > STACK) {
            overflow = true;
            break;
        }
        if (__builtin_amdgcn_inverse_ballot_w64(pb)) {
            // later lanes end lower, so lane 0's children are popped first
            const int slot = sp + npush - 1 - mbcnt(pb);
            stkRange[slot] = (uint32_t)tfirst | ((uint32_t)(tnchild - 1) << 28);
            stkMask[slot] = omask;
        }
        sp += npush;
        maxsp = max(maxsp, sp);
        __syncthreads();
    }

    const unsigned long long gpp = wave_sum_u64(npp);
    const unsigned long long gpc = wave_sum_u64(npc);
    nps = wave_sum_i(nps);
    const bool listOverflow = nf > lists.cf || np > lists.cp || nlf + nlp > lists.cl;
    if (lane == 0) {
        GroupCounts c;
        c.nf = nf; c.np = np; c.nlf = nlf; c.nlp = nlp; c.nps = nps;
        c.overflow = overflow ? 1 : 0;
        c.pad[0] = c.pad[1] = 0;
        lists.counts[r] = c;
        atomicAdd(&meta->pp, gpp);
        atomicAdd(&meta->pc, gpc);
        atomicMax(&meta->maxStack, maxsp);
        atomicMax(&meta->maxList[0], nf);
        atomicMax(&meta->maxList[1], np);
        atomicMax(&meta->maxList[2], nlf + nlp);
        atomicAdd(&meta->listEntries, (unsigned long long)(nf + np));
        atomicAdd(&meta->rounds, (unsigned long long)nrounds);
        atomicAdd(&meta->mixed, (unsigned long long)nmixed);
        if (overflow) atomicOr(&meta->flags, (int)kFlagStackOverflow);
        if (listOverflow) atomicOr(&meta->flags, (int)kFlagListOverflow);
    }
    return !overflow && !listOverflow;
}

// ---------------------------------------------------------------------------
// evaluation: one wavefront per group, lists from HBM
// ---------------------------------------------------------------------------
struct Sums {
    float sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;  // Kahan sums
    __device__ void add(float ax, float ay, float az) {
        kahan_add(sx, cx, ax);
        kahan_add(sy, cy, ay);
        kahan_add(sz, cz, az);
    }
};

// Stars of up to 64 leaf ranges (one per lane: lcur, lcnt stars, mask lmask)
// into the staging buffer, which is evaluated (flush) whenever it is full.
// All lanes copy: star j of the batch goes to lane j % 64, which finds the
// range holding it by a binary search over the ranges' exclusive prefix sums
// (tables beside the staged stars).  MASKED: the staged stars refer to this
// call's table of masks, so they are evaluated before the call returns.
template <bool MASKED, int PARTS, typename Flush>
__device__ __forceinline__ void stage_leaves(EvalLDS<PARTS>& S, int& nst, int lcur, int lcnt, uint64_t lmask,
                                             const double4* __restrict__ spos, const Group& G, Flush flush) {
    int* tStart = S.s.tStart;
    int* tCur = S.s.tCur;
    uint64_t* tMask = S.s.tMask;
    const int lane = lane_id();
    const int incl = wave_incl_scan(lcnt);
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, kWave));
    if (total == 0) return;
    tStart[lane] = incl - lcnt;
    tCur[lane] = lcur;
    if (MASKED) tMask[lane] = lmask;
    __syncthreads();
    for (int done = 0; done < total;) {
        const int take = min(PARTS - nst, total - done);
        for (int j0 = 0; j0 < take; j0 += kWave) {
            if (j0 + lane < take) {
                const int j = done + j0 + lane;
                // the last range starting at or before j (empty ranges share
                // the next one's start and are passed over)
                int o = 0;
#pragma unroll
                for (int step = kWave / 2; step > 0; step >>= 1)
                    if (tStart[o + step] <= j) o += step;
                const double4 q = spos[tCur[o] + (j - tStart[o])];
                const int at = nst + j0 + lane;
                S.s.part[at] = make_float4((float)(q.x - G.cgx), (float)(q.y - G.cgy), (float)(q.z - G.cgz),
                                         (float)q.w);
                if (MASKED) S.s.range[at] = (uint8_t)o;
            }
        }
        nst += take;
        done += take;
        if (nst == PARTS) flush();
    }
    if (MASKED && nst > 0) flush();
    __syncthreads();  // the tables are rewritten by the next batch
}

template <bool QUAD, bool EPS0, bool Q16, int PARTS, bool PREFETCH>
__device__ void eval_group(const Group& G, long long r, EvalLDS<PARTS>& S, const double4* __restrict__ spos,
                           const TravNode* __restrict__ trav, const Mult32* __restrict__ m32,
                           const Mult16* __restrict__ m16, const int32_t* __restrict__ order, const Lists& lists,
                           double* __restrict__ oax, double* __restrict__ oay, double* __restrict__ oaz) {
    const float eps2 = G.eps2;
    float4* cell = S.cell;
    const int lane = lane_id();
    const float px = (float)(G.p.x - G.cgx), py = (float)(G.p.y - G.cgy), pz = (float)(G.p.z - G.cgz);
    const GroupCounts cnts = lists.counts[r];
    // list lengths are uniform over the wave: they live in SGPRs
    const int nfc = __builtin_amdgcn_readfirstlane(cnts.nf);
    const int npc = __builtin_amdgcn_readfirstlane(cnts.np);
    const int nlf = __builtin_amdgcn_readfirstlane(cnts.nlf);
    const int nlp = __builtin_amdgcn_readfirstlane(cnts.nlp);
    const int nps = __builtin_amdgcn_readfirstlane(cnts.nps);
    const uint32_t* __restrict__ fullIn = lists.full + (size_t)r * lists.cf;
    const uint3* __restrict__ partIn = lists.part + (size_t)r * lists.cp;
    const uint3* __restrict__ leafIn = lists.leaf + (size_t)r * lists.cl;
    Sums sum;

    // With PREFETCH the cell chunks are software-pipelined: while chunk i is
    // evaluated from LDS, the node data of chunk i+1 and the indices of chunk
    // i+2 are in flight.
    struct Staged {
        float4 a, b;
        float2 c;
        uint64_t m;
    };
    auto fetch = [&](int node, uint64_t m) {
        Staged d;
        const TravNode& t = trav[node];
        if (Q16) {
            const Mult16 mh = m16[node];
            const float h = (float)ldexp(G.half, -t.level);
            const float sc = mh.m * h * h;
            auto q = [&](int c) { return __half2float(__ushort_as_half(mh.q[c])) * sc; };
            d.a = make_float4((float)(t.x - G.cgx), (float)(t.y - G.cgy), (float)(t.z - G.cgz), mh.m);
            d.b = make_float4(q(0), q(1), q(2), q(3));
            d.c = make_float2(q(4), q(5));
        } else {
            const Mult32 mm = m32[node];
            d.a = make_float4((float)(t.x - G.cgx), (float)(t.y - G.cgy), (float)(t.z - G.cgz), mm.m);
            d.b = make_float4(mm.qxx, mm.qyy, mm.qzz, mm.qxy);
            d.c = make_float2(mm.qxz, mm.qyz);
        }
        d.m = m;
        return d;
    };
    auto index_of = [&](bool partial, int i, uint64_t& m) -> int {
        if (partial) {
            const uint3 e = at32(partIn, i);
            m = (uint64_t)e.x | ((uint64_t)e.y << 32);
            return (int)e.z;
        }
        m = G.gmask;
        return (int)at32(fullIn, i);
    };
    // Full cells go in list order, chunk c holding entries [64c, 64c + 64).
    // Partial cells are dealt to the chunks with a stride (chunk c holds
    // entries c, c + nch, c + 2 nch, ...).
    for (int pass = 0; pass < 2; pass++) {
        const bool partial = pass == 1;
        const int n = partial ? npc : nfc;
        if (n == 0) continue;
        const int nch = (n + kChunk - 1) / kChunk;
        auto entry = [&](int c, int j) { return partial ? c + nch * j : c * kChunk + j; };
        if (!PREFETCH) {
            // at high occupancy the other wavefronts hide the fetch: no
            // registers held for the next chunk
            for (int c = 0; c < nch; c++) {
                const bool mine = entry(c, lane) < n;
                const int cnt = partial ? (n - c + nch - 1) / nch : min(kChunk, n - c * kChunk);
                uint64_t m = 0;
                if (mine) {
                    const int idx = index_of(partial, entry(c, lane), m);
                    const Staged d = fetch(idx, m);
                    cell[3 * lane] = d.a;
                    cell[3 * lane + 1] = d.b;
                    cell[3 * lane + 2] = make_float4(d.c.x, d.c.y, 0.0f, 0.0f);
                }
                __syncthreads();
                float ax = 0, ay = 0, az = 0;
                if (partial)
                    eval_cells_own<QUAD>(cell, transpose_bits(mine ? m : 0ull), px, py, pz, eps2, ax, ay, az);
                else
                    eval_cells<QUAD>(cell, cnt, px, py, pz, eps2, ax, ay, az);
                sum.add(ax, ay, az);
                __syncthreads();
            }
            continue;
        }
        uint64_t m0 = 0, m1 = 0;
        int idx0 = entry(0, lane) < n ? index_of(partial, entry(0, lane), m0) : 0;
        int idx1 = nch > 1 && entry(1, lane) < n ? index_of(partial, entry(1, lane), m1) : 0;
        Staged cur;
        if (entry(0, lane) < n) cur = fetch(idx0, m0);
        for (int c = 0; c < nch; c++) {
            const int c0 = c * kChunk;
            const bool mine = entry(c, lane) < n;
            const int cnt = partial ? (n - c + nch - 1) / nch : min(kChunk, n - c0);
            if (mine) {
                cell[3 * lane] = cur.a;
                cell[3 * lane + 1] = cur.b;
                cell[3 * lane + 2] = make_float4(cur.c.x, cur.c.y, 0.0f, 0.0f);
            }
            __syncthreads();
            Staged nxt;
            if (c + 1 < nch && entry(c + 1, lane) < n) nxt = fetch(idx1, m1);
            if (c + 2 < nch && entry(c + 2, lane) < n) idx1 = index_of(partial, entry(c + 2, lane), m1);
            float ax = 0, ay = 0, az = 0;
            if (partial)
                eval_cells_own<QUAD>(cell, transpose_bits(mine ? cur.m : 0ull), px, py, pz, eps2, ax, ay, az);
            else
                eval_cells<QUAD>(cell, cnt, px, py, pz, eps2, ax, ay, az);
            sum.add(ax, ay, az);
            __syncthreads();
            cur = nxt;
        }
    }

    // Stars of opened leaves (large leaves are split over flushes).
    int nst = 0;
    auto flush_all = [&]() {  // stars every lane uses
        __syncthreads();
        float ax = 0, ay = 0, az = 0;
        eval_parts<EPS0>(S.s.part, 0, nst, px, py, pz, eps2, ax, ay, az);
        sum.add(ax, ay, az);
        nst = 0;
        __syncthreads();
    };
    auto flush_masked = [&]() {  // stars of leaves some lanes opened: per-lane loops
        __syncthreads();
        float ax = 0, ay = 0, az = 0;
        eval_parts_own<EPS0>(S.s.part, S.s.range, S.s.tMask, 0, nst, px, py, pz, eps2, ax, ay, az);
        sum.add(ax, ay, az);
        nst = 0;
        __syncthreads();
    };
    // leaves opened by every star, in list order
    for (int c0 = 0; c0 < nlf; c0 += kWave) {
        int lcur = 0, lcnt = 0;
        if (c0 + lane < nlf) {
            const uint3 e = at32(leafIn, c0 + lane);
            lcur = trav[e.z].lo;
            lcnt = trav[e.z].hi - lcur;
        }
        stage_leaves<false, PARTS>(S, nst, lcur, lcnt, 0ull, spos, G, flush_all);
    }
    if (nst > 0) flush_all();
    // Leaves opened by some stars (stored from the back): flush f takes the
    // leaves f, f + nfl, f + 2 nfl, ..., so every flush mixes leaves from
    // all around the group and each lane's share of a flush's stars is close
    // to its share of the whole list.
    const int nfl = (nps + PARTS - 1) / PARTS;
    for (int f = 0; f < nfl; f++) {
        for (int c0 = 0; f + nfl * c0 < nlp; c0 += kWave) {
            const int j = f + nfl * (c0 + lane);
            int lcur = 0, lcnt = 0;
            uint64_t lmask = 0;
            if (j < nlp) {
                const uint3 e = at32(leafIn, lists.cl - 1 - j);
                lcur = trav[e.z].lo;
                lcnt = trav[e.z].hi - lcur;
                lmask = (uint64_t)e.x | ((uint64_t)e.y << 32);
            }
            stage_leaves<true, PARTS>(S, nst, lcur, lcnt, lmask, spos, G, flush_masked);
        }
    }
    if (G.valid) {
        const int32_t s = order[G.k];
        oax[s] = (double)sum.sx - (double)sum.cx;
        oay[s] = (double)sum.sy - (double)sum.cy;
        oaz[s] = (double)sum.sz - (double)sum.cz;
    }
}

// The two phases as separate kernels (GX_SPLIT_WALK=1, for timing them
// apart): every group keeps its own list region between the launches.
constexpr int kSplitStack = kWalkConfig[kWalkDefault].stack;
constexpr int kSplitParts = kWalkConfig[kWalkDefault].parts;

__global__ __launch_bounds__(64) void k_trav32(const double4* __restrict__ spos, const TravNode* __restrict__ trav,
                                               Meta* meta, Batch batch, Lists lists) {
    __shared__ TravLDS<kSplitStack> S;
    if (meta->flags & kFlagTreeIncomplete) return;
    Group G;
    const long long g = xcd_group(batch);
    if (!load_group(G, spos, batch, g)) return;
    trav_group<kSplitStack>(G, g, S, trav, meta, lists);
}

template <bool QUAD, bool EPS0, bool Q16>
__global__ __launch_bounds__(64) void k_eval32(const double4* __restrict__ spos, const TravNode* __restrict__ trav,
                                               const Mult32* __restrict__ m32, const Mult16* __restrict__ m16,
                                               const int32_t* __restrict__ order, Meta* meta, Batch batch,
                                               Lists lists, double* __restrict__ oax, double* __restrict__ oay,
                                               double* __restrict__ oaz) {
    __shared__ EvalLDS<kSplitParts> S;
    if (meta->flags & (kFlagTreeIncomplete | kFlagStackOverflow | kFlagListOverflow)) return;
    Group G;
    const long long g = xcd_group(batch);
    if (!load_group(G, spos, batch, g)) return;
    eval_group<QUAD, EPS0, Q16, kSplitParts, true>(G, g, S, spos, trav, m32, m16, order, lists, oax, oay, oaz);
}

// Traversal and evaluation of one group in one wavefront: the latency-bound
// traversal of some waves overlaps the VALU-bound evaluation of others on the
// same SIMD.  The lists still go through memory (they stay in L2/MALL between
// the two phases), so LDS holds only one phase's buffers at a time.  With
// persistent set, a grid of resident waves takes groups from a work queue
// (one launch for any number of groups and instances), and each workgroup
// reuses one list region (lists.perSlot).
template <bool QUAD, bool EPS0, bool Q16, int STACK, int PARTS, int WAVES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_walk32(const double4* __restrict__ spos, const TravNode* __restrict__ trav,
                                               const Mult32* __restrict__ m32, const Mult16* __restrict__ m16,
                                               const int32_t* __restrict__ order, Meta* meta, Batch batch,
                                               Lists lists, int persistent, double* __restrict__ oax,
                                               double* __restrict__ oay, double* __restrict__ oaz) {
    __shared__ WalkLDS<STACK, PARTS> S;
    if (meta->flags & kFlagTreeIncomplete) return;
    long long g = persistent ? -1 : xcd_group(batch);
    // The work queue is split in 8, one contiguous eighth of the groups per
    // XCD (workgroups are dealt round-robin to the XCDs): each XCD's L2 sees
    // the node records of one region of space.  An XCD that runs out helps
    // the next one.
    const long long ng = batch.groupOff[batch.B];
    const long long per = (ng + 7) / 8;
    int q = blockIdx.x % 8, tried = 0;
    for (;;) {
        if (persistent) {
            long long next = ng;
            if (lane_id() == 0)
                while (tried < 8) {
                    const long long t = (long long)q * per + (long long)atomicAdd(&meta->xcdNext[q], 1ull);
                    if (t < min(ng, (long long)(q + 1) * per)) {
                        next = t;
                        break;
                    }
                    q = (q + 1) % 8;
                    tried++;
                }
            // lane 0's claim, in scalar registers (the list regions' bases
            // and every per-group offset derived from it stay uniform)
            g = (long long)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(next >> 32)) << 32) |
                            (uint32_t)__builtin_amdgcn_readfirstlane((int)next));
        }
        Group G;
        if (!load_group(G, spos, batch, g)) return;
        const long long r = lists.perSlot ? (long long)blockIdx.x : g;
        const bool ok = trav_group<STACK>(G, r, S.t, trav, meta, lists);
        __syncthreads();  // lists written by this wave's lanes are read back by others
        if (ok) eval_group<QUAD, EPS0, Q16, PARTS, (WAVES < 6)>(G, r, S.e, spos, trav, m32, m16, order, lists, oax, oay, oaz);
        if (!persistent) return;
        __syncthreads();
    }
}

namespace {

using WalkKernel = void (*)(const double4*, const TravNode*, const Mult32*, const Mult16*, const int32_t*, Meta*,
                            Batch, Lists, int, double*, double*, double*);

// variant: 0 quadrupole, 1 quadrupole + FP16-stored, 2 monopole; x2 for eps0
template <int C>
constexpr WalkKernel walk_kernel(int variant) {
    constexpr int st = kWalkConfig[C].stack, pa = kWalkConfig[C].parts, wv = kWalkConfig[C].waves;
    switch (variant) {
        case 0: return k_walk32<true, false, false, st, pa, wv>;
        case 1: return k_walk32<true, false, true, st, pa, wv>;
        case 2: return k_walk32<false, false, false, st, pa, wv>;
        case 3: return k_walk32<true, true, false, st, pa, wv>;
        case 4: return k_walk32<true, true, true, st, pa, wv>;
        default: return k_walk32<false, true, false, st, pa, wv>;
    }
}

static_assert(kWalkConfigs == 9, "walk_kernel_of dispatches every configuration");
WalkKernel walk_kernel_of(int config, int variant) {
    switch (config) {
        case 0: return walk_kernel<0>(variant);
        case 1: return walk_kernel<1>(variant);
        case 2: return walk_kernel<2>(variant);
        case 3: return walk_kernel<3>(variant);
        case 4: return walk_kernel<4>(variant);
        case 5: return walk_kernel<5>(variant);
        case 6: return walk_kernel<6>(variant);
        case 7: return walk_kernel<7>(variant);
        default: return walk_kernel<8>(variant);
    }
}

}  // namespace

int walk32_blocks_per_cu(int config) {
    int per = 0;
    if (config < 0 || config >= kWalkConfigs) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)walk_kernel_of(config, 0), 64, 0) !=
        hipSuccess)
        return 0;
    return per;
}

void launch_walk32(const double4* spos, const TravNode* trav, const Mult32* m32, const Mult16* m16,
                   const int32_t* order, Meta* meta, const Batch& batch, long long ngroupsUpper, bool eps0,
                   int quad, const Lists& lists, double* ax, double* ay, double* az, hipStream_t stream,
                   hipEvent_t mid, bool split, int persistentBlocks, int config) {
    if (ngroupsUpper <= 0) return;
    const long long per = (ngroupsUpper + 7) / 8;
    const bool persistent = !split && persistentBlocks > 0;
    dim3 grid((unsigned)(persistent ? persistentBlocks : per * 8)), block(64);
    const bool q16 = m16 != nullptr;
    if (split) {
        hipLaunchKernelGGL(k_trav32, grid, block, 0, stream, spos, trav, meta, batch, lists);
        if (mid) (void)hipEventRecord(mid, stream);
#define GX_EVAL(Q, E, H) \
    hipLaunchKernelGGL((k_eval32<Q, E, H>), grid, block, 0, stream, spos, trav, m32, m16, order, meta, batch, lists, \
                       ax, ay, az)
        if (quad) {
            if (q16) {
                if (eps0) GX_EVAL(true, true, true); else GX_EVAL(true, false, true);
            } else {
                if (eps0) GX_EVAL(true, true, false); else GX_EVAL(true, false, false);
            }
        } else {
            if (eps0) GX_EVAL(false, true, false); else GX_EVAL(false, false, false);
        }
#undef GX_EVAL
        return;
    }
    const int variant = (quad ? (q16 ? 1 : 0) : 2) + (eps0 ? 3 : 0);
    hipLaunchKernelGGL(walk_kernel_of(config, variant), grid, block, 0, stream, spos, trav, m32, m16, order, meta,
                       batch, lists, persistent ? 1 : 0, ax, ay, az);
    if (mid) (void)hipEventRecord(mid, stream);
}

}  // namespace gx
