// gx_walk32.hip -- FP32 group walk for gfx950 (wave64).
//
// One wavefront walks the tree for a group of 64 consecutive stars in Morton
// order (one star per lane).  The wave shares one traversal stack and two
// interaction lists in LDS:
//   * cells     : accepted nodes, staged as (COM - c_g, m, Q) in FP32;
//   * particles : stars of opened leaves, staged as (x - c_g, m) in FP32,
// where c_g is the group's bounding-box centre.  Coordinates are made
// group-relative in FP64 before rounding, so FP32 only carries offsets that
// are small for near interactions.
//
// Every stack and list entry carries a 64-bit lane mask.  A node is tested
// against the group's bounding box in FP64; when the box does not decide it
// (some stars would accept the node and others open it), the wave evaluates
// the reference's own per-star test d2 >= crit2 in FP64 (bh.go:114-116) with a
// ballot, and the node is accepted for the lanes that pass and opened for the
// others.  The interaction set of every star is therefore exactly the one of
// the reference's per-particle walk (bh.go:106-155); only the arithmetic of
// the interactions is FP32.  Masks are applied with inverse ballots (the mask
// is used as the lane-select operand directly).
//
// Interactions follow bh.go:117-133 (cells, monopole + quadrupole) and
// bh.go:145-150 (stars, Plummer softening).  Each list flush is summed in
// FP32 and added to the per-star total with Kahan compensation.
#include "gx_common.h"
#include "gx_kernels.h"

namespace gx {

constexpr int kStack = 768;   // stack entries: (first child, child count, lane mask)
constexpr int kCells = 64;
constexpr int kParts = 256;
constexpr double kMargin = 1e-12;  // >> FP64 rounding of d2; box tests stay conservative

struct WalkLDS {
    float4 cell[kCells * 3];
    float4 part[kParts];
    uint64_t cellMask[kCells];
    uint64_t partMask[kParts];
    uint64_t stkMask[kStack];
    uint32_t stkRange[kStack];  // first | (count - 1) << 28
    uint64_t popMask[kWave];
    int popNode[kWave];
};

__device__ __forceinline__ double readlane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <bool QUAD>
__device__ __forceinline__ void eval_cells(const WalkLDS& L, int ncell, float px, float py, float pz, float eps2,
                                           float& ax, float& ay, float& az) {
    for (int e = 0; e < ncell; e++) {
        const uint64_t em = uniform64(L.cellMask[e]);
        const float4 c0 = L.cell[3 * e];
        const float dx = c0.x - px, dy = c0.y - py, dz = c0.z - pz;
        const float d2e = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        float ir = __builtin_amdgcn_rsqf(d2e);
        ir = __builtin_amdgcn_inverse_ballot_w64(em) ? ir : 0.0f;
        const float inv = ir * ir;
        const float inv3 = inv * ir;
        float f = c0.w * inv3;
        if (QUAD) {
            const float4 c1 = L.cell[3 * e + 1];
            const float4 c2 = L.cell[3 * e + 2];
            // Q = (qxx, qyy, qzz, qxy) in c1, (qxz, qyz) in c2.xy
            const float qx = fmaf(c1.x, dx, fmaf(c1.w, dy, c2.x * dz));
            const float qy = fmaf(c1.w, dx, fmaf(c1.y, dy, c2.y * dz));
            const float qz = fmaf(c2.x, dx, fmaf(c2.y, dy, c1.z * dz));
            const float inv5 = inv3 * inv;
            const float dqd = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
            f = fmaf(2.5f * dqd, inv5 * inv, f);
            ax = fmaf(-qx, inv5, ax);
            ay = fmaf(-qy, inv5, ay);
            az = fmaf(-qz, inv5, az);
        }
        ax = fmaf(f, dx, ax);
        ay = fmaf(f, dy, ay);
        az = fmaf(f, dz, az);
    }
}

template <bool EPS0>
__device__ __forceinline__ void eval_parts(const WalkLDS& L, int npart, float px, float py, float pz, float eps2,
                                           float& ax, float& ay, float& az) {
    for (int e = 0; e < npart; e++) {
        const uint64_t em = uniform64(L.partMask[e]);
        const float4 q = L.part[e];
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        float ir = __builtin_amdgcn_rsqf(r2);
        bool on = __builtin_amdgcn_inverse_ballot_w64(em);
        if (EPS0) on = on && r2 > 0.0f;  // the star itself: d = 0 exactly
        ir = on ? ir : 0.0f;
        const float mir = q.w * ir;
        const float f = mir * (ir * ir);
        ax = fmaf(f, dx, ax);
        ay = fmaf(f, dy, ay);
        az = fmaf(f, dz, az);
    }
}

template <bool QUAD, bool EPS0>
__global__ __launch_bounds__(64) void k_walk32(const double4* __restrict__ spos, const TravNode* __restrict__ trav,
                                               const Mult32* __restrict__ m32, const int32_t* __restrict__ order,
                                               Meta* meta, float eps2, double* __restrict__ oax,
                                               double* __restrict__ oay, double* __restrict__ oaz) {
    __shared__ WalkLDS L;
    const long long nst = meta->nstars;
    if (meta->flags & kFlagNodeOverflow) return;
    const long long ngroups = (nst + 63) / 64;
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so give
    // each XCD a contiguous run of (spatially adjacent) groups for its L2.
    const long long per = gridDim.x / 8;
    const long long g = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (g >= ngroups) return;
    const int lane = lane_id();
    const long long k = g * 64 + lane;
    const bool valid = k < nst;
    const double4 p = spos[valid ? k : g * 64];
    const uint64_t gmask = __ballot(valid);

    const double bminx = wave_min_d(valid ? p.x : INFINITY), bmaxx = wave_max_d(valid ? p.x : -INFINITY);
    const double bminy = wave_min_d(valid ? p.y : INFINITY), bmaxy = wave_max_d(valid ? p.y : -INFINITY);
    const double bminz = wave_min_d(valid ? p.z : INFINITY), bmaxz = wave_max_d(valid ? p.z : -INFINITY);
    const double cgx = 0.5 * (bminx + bmaxx), cgy = 0.5 * (bminy + bmaxy), cgz = 0.5 * (bminz + bmaxz);
    const float px = (float)(p.x - cgx), py = (float)(p.y - cgy), pz = (float)(p.z - cgz);

    float sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;  // Kahan sums
    unsigned long long npp = 0, npc = 0;
    int sp = 1, ncell = 0, npart = 0, maxsp = 1;
    bool overflow = false;
    if (lane == 0) {
        L.stkRange[0] = 0u;  // root, one node
        L.stkMask[0] = gmask;
    }
    __syncthreads();

    while (sp > 0) {
        // pop up to 64 nodes from the top ranges (top entry first)
        int cnt = 0;
        uint32_t rng = 0;
        uint64_t rmask = 0;
        if (lane < sp) {
            rng = L.stkRange[sp - 1 - lane];
            rmask = L.stkMask[sp - 1 - lane];
            cnt = (int)(rng >> 28) + 1;
        }
        const int incl0 = wave_incl_scan(cnt);
        const int excl0 = incl0 - cnt;
        const int take = max(0, min(cnt, kWave - excl0));
        const int rfirst = (int)(rng & 0x0fffffffu);
        for (int j = 0; j < take; j++) {
            L.popNode[excl0 + j] = rfirst + j;
            L.popMask[excl0 + j] = rmask;
        }
        const uint64_t full = __ballot(cnt > 0 && take == cnt);
        const int nfull = __popcll(full);
        const int npop = min(kWave, __shfl(incl0, 63, kWave));
        // a partly taken range stays on the stack with its remaining children
        if (lane == nfull && take > 0 && take < cnt)
            L.stkRange[sp - 1 - lane] = (uint32_t)(rfirst + take) | ((uint32_t)(cnt - take - 1) << 28);
        sp -= nfull;
        __syncthreads();
        const bool has = lane < npop;
        int node = 0;
        uint64_t mask = 0;
        if (has) {
            node = L.popNode[lane];
            mask = L.popMask[lane];
        }
        TravNode t;
        t.x = t.y = t.z = 0;
        t.crit2 = 0;
        t.first = -1;
        t.lo = t.hi = 0;
        t.nchild = 0;
        int cls = 0;  // 1 accept for all, 2 open for all, 3 decide per star
        if (has) {
            t = trav[node];
            const double ex = fmax(0.0, fmax(bminx - t.x, t.x - bmaxx));
            const double ey = fmax(0.0, fmax(bminy - t.y, t.y - bmaxy));
            const double ez = fmax(0.0, fmax(bminz - t.z, t.z - bmaxz));
            const double dmin2 = ex * ex + ey * ey + ez * ez;
            const double fx = fmax(fabs(t.x - bminx), fabs(t.x - bmaxx));
            const double fy = fmax(fabs(t.y - bminy), fabs(t.y - bmaxy));
            const double fz = fmax(fabs(t.z - bminz), fabs(t.z - bmaxz));
            const double dmax2 = fx * fx + fy * fy + fz * fz;
            if (dmin2 >= t.crit2 * (1 + kMargin))
                cls = 1;
            else if (dmax2 * (1 + kMargin) < t.crit2)
                cls = 2;
            else
                cls = 3;
        }
        uint64_t amask = cls == 1 ? mask : 0;
        uint64_t omask = cls == 2 ? mask : 0;
        uint64_t mixed = __ballot(cls == 3);
        while (mixed) {
            const int src = __builtin_ctzll(mixed);
            mixed &= mixed - 1;
            const double nx = readlane_d(t.x, src), ny = readlane_d(t.y, src), nz = readlane_d(t.z, src);
            const double ncrit2 = readlane_d(t.crit2, src);
            const uint64_t smask = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(mask >> 32), src) << 32) |
                                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mask, src);
            // the reference's per-star opening test, bh.go:114-116 (FP64, unfused)
            const double dx = nx - p.x, dy = ny - p.y, dz = nz - p.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const uint64_t acc = __ballot(d2 >= ncrit2) & smask;
            if (lane == src) {
                amask = acc;
                omask = smask & ~acc;
            }
        }

        // accepted cells -> cell list
        const bool isCell = amask != 0;
        const uint64_t cb = __ballot(isCell);
        const int nc = __popcll(cb);
        if (ncell + nc > kCells) {
            __syncthreads();
            float ax = 0, ay = 0, az = 0;
            eval_cells<QUAD>(L, ncell, px, py, pz, eps2, ax, ay, az);
            kahan_add(sx, cx, ax);
            kahan_add(sy, cy, ay);
            kahan_add(sz, cz, az);
            ncell = 0;
            __syncthreads();
        }
        if (isCell) {
            const int pos = ncell + mbcnt(cb);
            const Mult32 mm = m32[node];
            L.cell[3 * pos] = make_float4((float)(t.x - cgx), (float)(t.y - cgy), (float)(t.z - cgz), mm.m);
            L.cell[3 * pos + 1] = make_float4(mm.qxx, mm.qyy, mm.qzz, mm.qxy);
            L.cell[3 * pos + 2] = make_float4(mm.qxz, mm.qyz, 0.0f, 0.0f);
            L.cellMask[pos] = amask;
            npc += (unsigned long long)__popcll(amask);
        }
        ncell += nc;

        // opened internal nodes -> one stack entry with the whole child range
        const bool openInt = omask != 0 && t.first >= 0;
        const uint64_t pb = __ballot(openInt);
        const int npush = __popcll(pb);
        if (sp + npush > kStack) {
            overflow = true;
            break;
        }
        if (openInt) {
            // later lanes end lower, so lane 0's children are popped first
            const int slot = sp + npush - 1 - mbcnt(pb);
            L.stkRange[slot] = (uint32_t)t.first | ((uint32_t)(t.nchild - 1) << 28);
            L.stkMask[slot] = omask;
        }
        sp += npush;
        maxsp = max(maxsp, sp);

        // opened leaves -> particle list (large leaves are split over flushes)
        const bool openLeaf = omask != 0 && t.first < 0;
        int lcur = openLeaf ? t.lo : 0;
        int lcnt = openLeaf ? t.hi - t.lo : 0;
        if (openLeaf) npp += (unsigned long long)__popcll(omask) * (unsigned long long)lcnt;
        while (__ballot(lcnt > 0)) {
            const int room = kParts - npart;
            const int incl2 = wave_incl_scan(lcnt);
            const int excl2 = incl2 - lcnt;
            const int wcount = max(0, min(lcnt, room - excl2));
            for (int i = 0; i < wcount; i++) {
                const double4 q = spos[lcur + i];
                const int slot = npart + excl2 + i;
                L.part[slot] = make_float4((float)(q.x - cgx), (float)(q.y - cgy), (float)(q.z - cgz), (float)q.w);
                L.partMask[slot] = omask;
            }
            npart += wave_sum_i(wcount);
            lcur += wcount;
            lcnt -= wcount;
            if (__ballot(lcnt > 0)) {
                __syncthreads();
                float ax = 0, ay = 0, az = 0;
                eval_parts<EPS0>(L, npart, px, py, pz, eps2, ax, ay, az);
                kahan_add(sx, cx, ax);
                kahan_add(sy, cy, ay);
                kahan_add(sz, cz, az);
                npart = 0;
                __syncthreads();
            }
        }
        __syncthreads();
    }

    __syncthreads();
    {
        float ax = 0, ay = 0, az = 0;
        eval_cells<QUAD>(L, ncell, px, py, pz, eps2, ax, ay, az);
        eval_parts<EPS0>(L, npart, px, py, pz, eps2, ax, ay, az);
        kahan_add(sx, cx, ax);
        kahan_add(sy, cy, ay);
        kahan_add(sz, cz, az);
    }
    if (valid) {
        const int32_t s = order[k];
        oax[s] = overflow ? NAN : (double)sx - (double)cx;
        oay[s] = overflow ? NAN : (double)sy - (double)cy;
        oaz[s] = overflow ? NAN : (double)sz - (double)cz;
    }
    npp = wave_sum_u64(npp);
    npc = wave_sum_u64(npc);
    if (lane == 0) {
        atomicAdd(&meta->pp, npp);
        atomicAdd(&meta->pc, npc);
        atomicMax(&meta->maxStack, maxsp);
        if (overflow) atomicOr(&meta->flags, (int)kFlagStackOverflow);
    }
}

void launch_walk32(const double4* spos, const TravNode* trav, const Mult32* m32, const int32_t* order, Meta* meta,
                   long long ngroupsUpper, float eps2, int quad, double* ax, double* ay, double* az,
                   hipStream_t stream) {
    if (ngroupsUpper <= 0) return;
    const long long per = (ngroupsUpper + 7) / 8;
    dim3 grid((unsigned)(per * 8)), block(64);
    const bool eps0 = !(eps2 > 0.0f);
    if (quad) {
        if (eps0)
            hipLaunchKernelGGL((k_walk32<true, true>), grid, block, 0, stream, spos, trav, m32, order, meta, eps2,
                               ax, ay, az);
        else
            hipLaunchKernelGGL((k_walk32<true, false>), grid, block, 0, stream, spos, trav, m32, order, meta, eps2,
                               ax, ay, az);
    } else {
        if (eps0)
            hipLaunchKernelGGL((k_walk32<false, true>), grid, block, 0, stream, spos, trav, m32, order, meta, eps2,
                               ax, ay, az);
        else
            hipLaunchKernelGGL((k_walk32<false, false>), grid, block, 0, stream, spos, trav, m32, order, meta, eps2,
                               ax, ay, az);
    }
}

}  // namespace gx
