// gx_drag.hip -- frame dragging of spinning black holes on the GPU,
// internal/sim/framedrag.go.
//
// The reference kicks bodies one after another in a fixed order (dragOuter:
// state order; dragFast: the near list, then the holes).  Each kick rotates
// the body's velocity in the gravitomagnetic field of the holes' spins and
// books the opposite momentum and angular momentum on the holes, so every kick
// reads the spins that all earlier kicks left.  The pass over a list of
// bodies is evaluated here as a fixed point that is bit-identical to that
// sequential loop:
//   1. every body j of the list is kicked in parallel with a guess of the
//      spins it sees (first the spins at the start of the pass, then the
//      previous iterate), giving its spin increments dJ_j, one per hole and
//      component (-0.0 where the body gives nothing: x + -0.0 == x for every
//      x, signed zeros included);
//   2. the next guess is the exclusive prefix sums of the dJ in list order:
//      approximate in the first two iterations, then rounded exactly as the
//      sequential loop rounds them (the exact ordered scan below);
//   3. if an exact scan equals the guess bit for bit, the guess is a fixed
//      point of the sequential recursion S_{j+1} = S_j + dJ(S_j), S_0 = the
//      start spins, whose solution is unique, so it is the reference's chain.
//      Each exact iteration is exact at least one position further than the
//      one before; a kick changes a spin by a tiny relative amount, so in
//      practice the first or second exact iteration is the fixed point.  Past
//      an iteration limit the pass runs as one sequential thread.
// The holes' velocities are sums of the same kind: once the spins are exact,
// one more ordered scan of their increments.  A body's own velocity depends
// only on the spins it sees.
//
// Exact ordered scan.  A chain s_{t+1} = fl(s_t + d_t) is cut into segments of
// kDragSeg elements.  Every segment is first added element by element, all at
// once, from a guess of its start (the previous iterate's value there); a
// walk over the segments in order takes a segment from there when the guess
// is its exact start bit for bit, which in the iteration that confirms the
// fixed point is every segment.  Otherwise: while s_t lies strictly inside
// one binade, |s_t| in (2^E, 2^(E+1)), a step is s_{t+1} = s_t + u rint(d_t /
// u), u = 2^(E-52), whenever d_t / u is not a half-integer: the exact sum
// lies within u/2 of that grid point, which is inside the binade, and the
// grid there is u.  A half-integer (a tie) rounds to the even grid point,
// which depends on the running value only through its parity; after a tie
// the value is even, so only the first tie of a segment depends on its start
// (one bit, carried symbolically).  So a segment whose partial sums all stay
// inside the binade of its start value is an integer sum, associative: its
// total and its partial minimum and maximum are computed in parallel (one
// wavefront per segment), and the walk (one thread per chain, from the exact
// start value) checks the binade condition from those extremes.  A
// segment that fails it (a binade crossing, a zero or subnormal value, an
// element as large as the sum, a wrong binade guess) is redone by one
// wavefront in blocks of 64 from its exact start value, with the same
// integer rule and the binade read from that value; only a block that leaves
// its binade is added element by element, as the reference adds.  The
// binade of a segment is guessed from an approximate prefix sum.
//
// Compiled with -ffp-contract=off: the kick repeats framedrag.go's
// operations in its order.  One deliberate difference: dragBack's weights
// |J| / r^3 use r2 * sqrt(r2) for math.Pow(r2, 1.5), which Go evaluates
// through Exp and Log (a few ulp apart); the weights only enter as
// w / sum(w), which is exactly 1 with one spinning hole.
#include "gx_common.h"
#include "gx_kernels.h"
#include "gx_pair.h"

namespace gx {

namespace {

__device__ __forceinline__ double pow15(double x) { return x * sqrt(x); }

__device__ __forceinline__ unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }

constexpr unsigned long long kNegZero = 0x8000000000000000ull;

// Spins seen by the body at list position j: the guess S (chain-major,
// S[(3k + comp) * ne + j]) or, with S null, the holes' spins.
struct SpinView {
    const double* S;
    long long ne, j;
    const double *jx, *jy, *jz;
    const int32_t* holes;
    __device__ __forceinline__ void operator()(int k, double& x, double& y, double& z) const {
        if (S) {
            x = S[(3LL * k) * ne + j];
            y = S[(3LL * k + 1) * ne + j];
            z = S[(3LL * k + 2) * ne + j];
        } else {
            const int32_t h = holes[k];
            x = jx[h];
            y = jy[h];
            z = jz[h];
        }
    }
};

// The spins in memory (the sequential kernels).
struct SpinLive {
    const double *jx, *jy, *jz;
    const int32_t* holes;
    __device__ __forceinline__ void operator()(int k, double& x, double& y, double& z) const {
        const int32_t h = holes[k];
        x = jx[h];
        y = jy[h];
        z = jz[h];
    }
};

// gravitomagnetic (framedrag.go:34-54) and the Boris rotation of dragKick
// (:56-72) of body i with proper velocity u; false if B = 0 (no kick).
template <class Spin>
__device__ __forceinline__ bool drag_rotate(long long i, const DragArgs& a, const Spin& spin, double ux, double uy,
                                            double uz, double& unx, double& uny, double& unz) {
    const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
    double bx = 0, by = 0, bz = 0;
    for (int k = 0; k < a.nh; k++) {
        const int32_t h = a.holes[k];
        if (h == i || !a.alive[h]) continue;
        double jx, jy, jz;
        spin(k, jx, jy, jz);
        if (jx * jx + jy * jy + jz * jz == 0) continue;
        const double rx = xi - a.x[h], ry = yi - a.y[h], rz = zi - a.z[h];
        const double r2 = (rx * rx + ry * ry + rz * rz) + a.eps2;
        const double inv = 1 / sqrt(r2);
        const double k3 = 2 / (a.c * a.c) * inv * inv * inv;
        const double sc = 3 * (jx * rx + jy * ry + jz * rz) / r2;
        bx = bx + (jx - rx * sc) * k3;
        by = by + (jy - ry * sc) * k3;
        bz = bz + (jz - rz * sc) * k3;
    }
    if (bx * bx + by * by + bz * bz == 0) return false;
    const double g = gamma_of(ux, uy, uz, a.c);
    const double hs = 0.5 * a.dt / g;
    const double tx = bx * hs, ty = by * hs, tz = bz * hs;
    const double upx = ux + (uy * tz - uz * ty), upy = uy + (uz * tx - ux * tz), upz = uz + (ux * ty - uy * tx);
    const double s2 = 2 / (1 + (tx * tx + ty * ty + tz * tz));
    const double sx = tx * s2, sy = ty * s2, sz = tz * s2;
    unx = ux + (upy * sz - upz * sy);
    uny = uy + (upz * sx - upx * sz);
    unz = uz + (upx * sy - upy * sx);
    return true;
}

// dragBack's weight of hole slot k for body i (framedrag.go:82-93).
template <class Spin>
__device__ __forceinline__ double drag_weight(long long i, int k, const DragArgs& a, const Spin& spin) {
    const int32_t h = a.holes[k];
    if (h == i || !a.alive[h]) return 0;
    const double rx = a.x[i] - a.x[h], ry = a.y[i] - a.y[h], rz = a.z[i] - a.z[h];
    double jx, jy, jz;
    spin(k, jx, jy, jz);
    const double j = sqrt(jx * jx + jy * jy + jz * jz);
    if (j > 0) return j / pow15(rx * rx + ry * ry + rz * rz);
    return 0;
}

template <class Spin>
__device__ __forceinline__ double drag_total(long long i, const DragArgs& a, const Spin& spin) {
    double total = 0;
    for (int k = 0; k < a.nh; k++) total += drag_weight(i, k, a, spin);
    return total;
}

// One dragKick (framedrag.go:56-113) on the arrays in memory, for the
// sequential kernels.
__device__ void drag_kick_live(long long i, const DragArgs& a, double* vx, double* vy, double* vz, double* jx,
                               double* jy, double* jz) {
    const SpinLive spin{jx, jy, jz, a.holes};
    const double ux = vx[i], uy = vy[i], uz = vz[i];
    double unx, uny, unz;
    if (!drag_rotate(i, a, spin, ux, uy, uz, unx, uny, unz)) return;
    vx[i] = unx;
    vy[i] = uny;
    vz[i] = unz;
    const double mi = a.mass[i];
    const double pxm = (unx - ux) * mi, pym = (uny - uy) * mi, pzm = (unz - uz) * mi;
    const double total = drag_total(i, a, spin);
    if (total == 0) return;
    for (int k = 0; k < a.nh; k++) {
        const double w = drag_weight(i, k, a, spin);
        if (w == 0) continue;
        const int32_t h = a.holes[k];
        const double f = w / total, mh = a.mass[h];
        vx[h] += (pxm * -f) / mh;
        vy[h] += (pym * -f) / mh;
        vz[h] += (pzm * -f) / mh;
        const double rx = a.x[i] - a.x[h], ry = a.y[i] - a.y[h], rz = a.z[i] - a.z[h];
        const double gx = pxm * f, gy = pym * f, gz = pzm * f;
        jx[h] += (ry * gz - rz * gy) * -1;
        jy[h] += (rz * gx - rx * gz) * -1;
        jz[h] += (rx * gy - ry * gx) * -1;
    }
}

// A 64-bit value moved across lanes by one DPP control (both halves); lanes
// with no source, or in rows outside RowMask, receive 0.
template <int Ctrl, int RowMask, bool ZeroOutside>
__device__ __forceinline__ long long dpp64(long long v) {
    const int lo = (int)(unsigned)(unsigned long long)v, hi = (int)((unsigned long long)v >> 32);
    const unsigned l = (unsigned)__builtin_amdgcn_update_dpp(0, lo, Ctrl, RowMask, 0xf, ZeroOutside);
    const unsigned h = (unsigned)__builtin_amdgcn_update_dpp(0, hi, Ctrl, RowMask, 0xf, ZeroOutside);
    return (long long)(((unsigned long long)h << 32) | l);
}

// Inclusive prefix sum over the 64 lanes of a wavefront (all lanes active):
// row_shr 1, 2, 4, 8 within rows of 16, then the last lane of row 0 into row
// 1 and of row 2 into row 3 (row_bcast:15), and of row 1 into rows 2 and 3
// (row_bcast:31).  VALU only, no LDS.
__device__ __forceinline__ long long wave_incl_scan(long long v) {
    v += dpp64<0x111, 0xf, true>(v);
    v += dpp64<0x112, 0xf, true>(v);
    v += dpp64<0x114, 0xf, true>(v);
    v += dpp64<0x118, 0xf, true>(v);
    v += dpp64<0x142, 0xa, false>(v);
    v += dpp64<0x143, 0xc, false>(v);
    return v;
}

// The integer step of element d in binade e (see the header), q = d / u with
// u = 2^(e-52): big if |q| >= 2^53 (or not finite); else a tie if q is a
// half-integer (the exact sum lies halfway between two grid points and rounds
// to the one with the even integer), with base = floor(q), or base = rint(q).
// ldexp by a power of two is exact unless the result is subnormal, and then
// |q| < 2^-1022: base 0, no tie either way.
__device__ __forceinline__ long long seg_step(double d, int e, bool& big, bool& tie) {
    const double q = ldexp(d, 52 - e);
    big = !(fabs(q) < 9007199254740992.0);
    if (big) {
        tie = false;
        return 0;
    }
    tie = fabs(q - trunc(q)) == 0.5;
    return (long long)(tie ? floor(q) : rint(q));
}

__device__ __forceinline__ int dexp(double v) { return (int)((dbits(v) >> 52) & 0x7ff); }

// Integer scan along a chain with ties, one wavefront, row after row of 64
// elements.  A tie at element j adds base_j + a_j, a_j the parity of the
// integer just below the halfway point (round half to even).  Every tie
// leaves an even value, so a_j of each tie after the first is the parity of
// the bases added since the previous tie; the first tie's a_f is the parity
// of M + Q_f (M the start, Q_f the bases up to and including that tie) --
// used when M is known, else taken as 0 here and added by the caller (the
// segment summaries, computed before M is).
struct TieState {
    long long Q = 0, A = 0;            // bases and adjustments of the rows so far
    long long lastQ = 0, firstQ = 0;   // Q at the last and at the first tie
    bool seen = false;                 // a tie so far
};

// One row: lane l's base and tie flag (0 and false past the data).  Returns
// the inclusive partial sum (bases + adjustments, relative to M); own = this
// lane's base + adjustment; after = a tie lies at or before this lane.
__device__ __forceinline__ long long tie_row(long long base, bool tie, bool knownM, long long M, TieState& st,
                                             long long& own, bool& after) {
    const int lane = threadIdx.x & 63;
    const long long Q = st.Q + wave_incl_scan(base);
    const unsigned long long tm = __ballot(tie);
    const unsigned long long below = tm & ((1ULL << lane) - 1);
    const long long Qp = __shfl(Q, below ? 63 - __clzll((long long)below) : lane, 64);
    long long a = 0;
    if (tie) {
        if (below)
            a = (Q - Qp) & 1;
        else if (st.seen)
            a = (Q - st.lastQ) & 1;
        else
            a = knownM ? ((M + Q) & 1) : 0;
    }
    const long long A = st.A + wave_incl_scan(a);
    own = base + a;
    after = st.seen || (tm & ((2ULL << lane) - 1)) != 0;
    const long long qEnd = __shfl(Q, 63, 64), aEnd = __shfl(A, 63, 64);
    if (tm) {
        const long long ql = __shfl(Q, 63 - __clzll((long long)tm), 64);
        const long long qf = __shfl(Q, __ffsll((long long)tm) - 1, 64);
        if (!st.seen) st.firstQ = qf;
        st.lastQ = ql;
        st.seen = true;
    }
    st.Q = qEnd;
    st.A = aEnd;
    return Q + A;
}

// One block of up to 64 elements of a chain, by one wavefront, from the exact
// value v (wave-uniform; lane l holds element l in d, -0.0 past len): as an
// integer prefix sum in the binade of v when every partial value stays
// strictly inside it (the exactness argument of the header, with the binade
// read from v itself and ties rounded to even), else element by element,
// every lane carrying the same running value and reading element t from
// lane t.  out[l] (may be null) receives the value before element l; returns
// the value after the block.
__device__ double chain_block(double v, double d, int len, double* out) {
    const int lane = threadIdx.x & 63;
    const long long two52 = 1LL << 52, two53 = 1LL << 53;
    const int ev = dexp(v);
    if (ev != 0 && ev != 0x7ff) {
        const int e = ev - 1023;
        const long long M = (long long)ldexp(v, 52 - e);  // exact: |M| in [2^52, 2^53)
        bool big, tie;
        const long long n = seg_step(d, e, big, tie);
        TieState st;
        long long own;
        bool after;
        const long long P = M + tie_row(n, tie, true, M, st, own, after);  // |P| < 2^53 + 65 * 2^53
        const bool in = M > 0 ? (P > two52 && P < two53) : (P > -two53 && P < -two52);
        const bool mok = M > 0 ? M > two52 : M < -two52;
        if (mok && !__any(big || !in)) {
            if (out && lane < len) out[lane] = ldexp((double)(P - own), e - 52);
            return ldexp((double)__shfl(P, 63, 64), e - 52);
        }
    }
    const unsigned long long db = dbits(d);
    const int dlo = (int)(unsigned)db, dhi = (int)(db >> 32);
    double r = v, mine = 0;
    for (int t = 0; t < len; t++) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane(dlo, t);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane(dhi, t);
        const double x = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        mine = lane == t ? r : mine;
        r = r + x;
    }
    if (out && lane < len) out[lane] = mine;
    return r;
}

}  // namespace

__global__ void k_drag_spinning(const int32_t* __restrict__ holes, const Meta* __restrict__ meta,
                                const double* __restrict__ jx, const double* __restrict__ jy,
                                const double* __restrict__ jz, double c, int* __restrict__ on) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int v = 0;
    if (c > 0)
        for (long long k = 0; k < meta->nholes; k++) {
            const int32_t h = holes[k];
            if (jx[h] != 0 || jy[h] != 0 || jz[h] != 0) v = 1;
        }
    *on = v;
}

// dragOuter's bodies (framedrag.go:115-127): alive, not a hole, not near.
__global__ __launch_bounds__(256) void k_drag_outer_flag(long long n, const uint8_t* __restrict__ kind,
                                                         const uint8_t* __restrict__ alive,
                                                         const uint8_t* __restrict__ nearFlag,
                                                         uint8_t* __restrict__ flag) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        flag[i] = kind[i] != kHole && alive[i] && !(nearFlag && nearFlag[i]);
}

// The chain starts: spins (mode 0) or velocities (mode 1) of the holes,
// start[3k + comp]; with mode 2 the chain ends are written back instead.
__global__ void k_drag_chain_io(const int32_t* __restrict__ holes, int nh, double* __restrict__ ax,
                                double* __restrict__ ay, double* __restrict__ az, double* __restrict__ vals,
                                int writeBack) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nh; k += gridDim.x * blockDim.x) {
        const int32_t h = holes[k];
        if (writeBack) {
            ax[h] = vals[3 * k];
            ay[h] = vals[3 * k + 1];
            az[h] = vals[3 * k + 2];
        } else {
            vals[3 * k] = ax[h];
            vals[3 * k + 1] = ay[h];
            vals[3 * k + 2] = az[h];
        }
    }
}

// Step 1 (apply = 0): the spin increments of every body of the list with the
// spins it sees (S, or the holes' spins with S null).  apply = 1: the body's
// new velocity is written and D receives the holes' velocity increments.
// Without frame dragging (*on == 0) every increment is -0.0.
__global__ __launch_bounds__(256) void k_drag_eval(DragArgs a, const int32_t* __restrict__ list, long long ne,
                                                   double* __restrict__ vx, double* __restrict__ vy,
                                                   double* __restrict__ vz, const double* __restrict__ jx,
                                                   const double* __restrict__ jy, const double* __restrict__ jz,
                                                   const double* __restrict__ S, int apply,
                                                   const int* __restrict__ on, double* __restrict__ D) {
    const bool active = *on != 0;
    const double nz = __longlong_as_double((long long)kNegZero);
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < ne; j += (long long)gridDim.x * blockDim.x) {
        const long long i = list[j];
        const SpinView spin{S, ne, j, jx, jy, jz, a.holes};
        const double ux = vx[i], uy = vy[i], uz = vz[i];
        double unx, uny, unz;
        const bool kicked = active && drag_rotate(i, a, spin, ux, uy, uz, unx, uny, unz);
        if (!kicked) {
            for (int k = 0; k < 3 * a.nh; k++) D[k * ne + j] = nz;
            continue;
        }
        if (apply) {
            vx[i] = unx;
            vy[i] = uny;
            vz[i] = unz;
        }
        const double mi = a.mass[i];
        const double pxm = (unx - ux) * mi, pym = (uny - uy) * mi, pzm = (unz - uz) * mi;
        const double total = drag_total(i, a, spin);
        for (int k = 0; k < a.nh; k++) {
            const double w = total == 0 ? 0.0 : drag_weight(i, k, a, spin);
            double o0 = nz, o1 = nz, o2 = nz;
            if (w != 0) {
                const int32_t h = a.holes[k];
                const double f = w / total;
                if (apply) {
                    const double mh = a.mass[h];
                    o0 = (pxm * -f) / mh;
                    o1 = (pym * -f) / mh;
                    o2 = (pzm * -f) / mh;
                } else {
                    const double rx = a.x[i] - a.x[h], ry = a.y[i] - a.y[h], rz = a.z[i] - a.z[h];
                    const double gx = pxm * f, gy = pym * f, gz = pzm * f;
                    o0 = (ry * gz - rz * gy) * -1;
                    o1 = (rz * gx - rx * gz) * -1;
                    o2 = (rx * gy - ry * gx) * -1;
                }
            }
            D[(3LL * k) * ne + j] = o0;
            D[(3LL * k + 1) * ne + j] = o1;
            D[(3LL * k + 2) * ne + j] = o2;
        }
    }
}

// The pass as one thread, in the reference's order: the fallback of the fixed
// point, and dragFast's holes (framedrag.go:138-142).
__global__ void k_drag_seq(DragArgs a, const int32_t* __restrict__ list, long long ne, double* vx, double* vy,
                           double* vz, double* jx, double* jy, double* jz, const int* __restrict__ on) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || !*on) return;
    for (long long j = 0; j < ne; j++) drag_kick_live(list[j], a, vx, vy, vz, jx, jy, jz);
}

// A whole pass over a short list in one workgroup, without the host (the
// near lists of dragFast, at most kDragSmall bodies, two passes per
// substep): the fixed point of the header with the chains added in order by
// one thread each (C = 3 nh threads; a short chain costs less than the scan's
// launches), the apply step, then, with holesToo, the holes one by one
// (framedrag.go:138-142); the sequential pass by thread 0 if it has not
// settled after maxIt iterations.  D, S0, S1: C * ne doubles each; E: 2 C
// doubles (chain ends).  counters[0] += iterations, counters[1] += 1 if
// sequential.
__global__ __launch_bounds__(1024) void k_drag_small(DragArgs a, const int32_t* __restrict__ list, long long ne,
                                                     double* vx, double* vy, double* vz, double* jx, double* jy,
                                                     double* jz, const int* __restrict__ on, double* D, double* S0,
                                                     double* S1, double* E, int maxIt, int holesToo,
                                                     unsigned long long* counters) {
    __shared__ int changed;
    const int tid = threadIdx.x, C = 3 * a.nh;
    if (!*on) return;
    const double nz = __longlong_as_double((long long)kNegZero);
    double* Sold = nullptr;
    double* Snew = S0;
    bool settled = false;
    int it = 0;
    auto eval = [&](const double* S, int apply) {
        for (long long j = tid; j < ne; j += blockDim.x) {
            const long long i = list[j];
            const SpinView spin{S, ne, j, jx, jy, jz, a.holes};
            const double ux = vx[i], uy = vy[i], uz = vz[i];
            double unx, uny, unz;
            if (!drag_rotate(i, a, spin, ux, uy, uz, unx, uny, unz)) {
                for (int k = 0; k < C; k++) D[k * ne + j] = nz;
                continue;
            }
            if (apply) {
                vx[i] = unx;
                vy[i] = uny;
                vz[i] = unz;
            }
            const double mi = a.mass[i];
            const double pxm = (unx - ux) * mi, pym = (uny - uy) * mi, pzm = (unz - uz) * mi;
            const double total = drag_total(i, a, spin);
            for (int k = 0; k < a.nh; k++) {
                const double w = total == 0 ? 0.0 : drag_weight(i, k, a, spin);
                double o0 = nz, o1 = nz, o2 = nz;
                if (w != 0) {
                    const int32_t h = a.holes[k];
                    const double f = w / total;
                    if (apply) {
                        const double mh = a.mass[h];
                        o0 = (pxm * -f) / mh;
                        o1 = (pym * -f) / mh;
                        o2 = (pzm * -f) / mh;
                    } else {
                        const double rx = a.x[i] - a.x[h], ry = a.y[i] - a.y[h], rz = a.z[i] - a.z[h];
                        const double gx = pxm * f, gy = pym * f, gz = pzm * f;
                        o0 = (ry * gz - rz * gy) * -1;
                        o1 = (rz * gx - rx * gz) * -1;
                        o2 = (rx * gy - ry * gx) * -1;
                    }
                }
                D[(3LL * k) * ne + j] = o0;
                D[(3LL * k + 1) * ne + j] = o1;
                D[(3LL * k + 2) * ne + j] = o2;
            }
        }
    };
    auto chain_start = [&](int c, const double* px, const double* py, const double* pz) {
        const int32_t h = a.holes[c / 3];
        return c % 3 == 0 ? px[h] : (c % 3 == 1 ? py[h] : pz[h]);
    };
    while (it < maxIt && !settled) {
        it++;
        eval(Sold, 0);
        if (tid == 0) changed = 0;
        __syncthreads();
        if (tid < C) {
            const double st = chain_start(tid, jx, jy, jz);
            double v = st;
            bool diff = false;
            const double* __restrict__ d = D + tid * ne;
            const double* __restrict__ so = Sold ? Sold + tid * ne : nullptr;
            double* __restrict__ sn = Snew + tid * ne;
            constexpr int kB = 16;  // loads issued ahead of the dependent adds
            for (long long j0 = 0; j0 < ne; j0 += kB) {
                double x[kB], o[kB];
#pragma unroll
                for (int k = 0; k < kB; k++) {
                    x[k] = j0 + k < ne ? d[j0 + k] : 0.0;
                    o[k] = so && j0 + k < ne ? so[j0 + k] : st;
                }
#pragma unroll
                for (int k = 0; k < kB; k++)
                    if (j0 + k < ne) {
                        diff = diff || dbits(o[k]) != dbits(v);
                        sn[j0 + k] = v;
                        v = v + x[k];
                    }
            }
            E[tid] = v;
            if (diff) changed = 1;
        }
        __syncthreads();
        settled = changed == 0;
        if (!settled) {
            Sold = Snew;
            Snew = Snew == S0 ? S1 : S0;
        }
        __syncthreads();
    }
    if (settled) {
        eval(Sold, 1);  // the bodies' velocities and the holes' velocity increments
        __syncthreads();
        if (tid < C) {
            double v = chain_start(tid, vx, vy, vz);
            const double* __restrict__ d = D + tid * ne;
            constexpr int kB = 16;
            for (long long j0 = 0; j0 < ne; j0 += kB) {
                double x[kB];
#pragma unroll
                for (int k = 0; k < kB; k++) x[k] = j0 + k < ne ? d[j0 + k] : 0.0;
#pragma unroll
                for (int k = 0; k < kB; k++)
                    if (j0 + k < ne) v = v + x[k];
            }
            E[C + tid] = v;
        }
        __syncthreads();
        if (tid < C) {
            const int32_t h = a.holes[tid / 3];
            double* pv = tid % 3 == 0 ? vx : (tid % 3 == 1 ? vy : vz);
            double* pj = tid % 3 == 0 ? jx : (tid % 3 == 1 ? jy : jz);
            pv[h] = E[C + tid];
            pj[h] = E[tid];
        }
        __syncthreads();
        if (tid == 0) atomicAdd(counters, (unsigned long long)it);
    } else if (tid == 0) {
        for (long long j = 0; j < ne; j++) drag_kick_live(list[j], a, vx, vy, vz, jx, jy, jz);
        atomicAdd(counters, (unsigned long long)it);
        atomicAdd(counters + 1, 1ULL);
    }
    __syncthreads();
    if (holesToo && tid == 0)
        for (int k = 0; k < a.nh; k++) drag_kick_live(a.holes[k], a, vx, vy, vz, jx, jy, jz);
}

// Exact ordered scan, step A: approximate sums of every segment and of its
// rows of kDragRow elements (one wavefront per segment; chain c, segment g).
__global__ __launch_bounds__(256) void k_scan_segsum(const double* __restrict__ D, long long ne, int chains, int nseg,
                                                     double* __restrict__ sig, double* __restrict__ rowSig) {
    const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= (long long)chains * nseg) return;
    const int c = (int)(w / nseg), g = (int)(w % nseg);
    const long long lo = (long long)g * kDragSeg, hi = min(ne, lo + kDragSeg);
    double v = 0;
    for (int r = 0; r < kDragRows; r++) {
        const long long t = lo + (long long)r * kDragRow + lane;
        double x = t < hi ? D[c * ne + t] : 0.0;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
        if (lane == 0) rowSig[w * kDragRows + r] = x;
        v += x;
    }
    if (lane == 0) sig[w] = v;
}

// Step B: approximate start values of the segments, guess[g] = start +
// sum of the segments before g (one workgroup of 1024 threads per chain).
__global__ __launch_bounds__(1024) void k_scan_guess(const double* __restrict__ sig, int nseg,
                                                     const double* __restrict__ start, double* __restrict__ guess) {
    __shared__ double buf[1024];
    const int c = blockIdx.x, tid = threadIdx.x;
    double carry = start[c];
    for (int base = 0; base < nseg; base += 1024) {
        const int g = base + tid;
        const double v = g < nseg ? sig[(long long)c * nseg + g] : 0.0;
        buf[tid] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const double o = tid >= d ? buf[tid - d] : 0.0;
            __syncthreads();
            buf[tid] += o;
            __syncthreads();
        }
        if (g < nseg) guess[(long long)c * nseg + g] = carry + (buf[tid] - v);
        carry += buf[1023];
        __syncthreads();
    }
}

// Step C: the integer summary of every segment in the binade of its guess.
__global__ __launch_bounds__(256) void k_scan_summary(const double* __restrict__ D, long long ne, int chains, int nseg,
                                                      const double* __restrict__ guess, DragSeg* __restrict__ seg) {
    const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= (long long)chains * nseg) return;
    const int c = (int)(w / nseg), g = (int)(w % nseg);
    const long long lo = (long long)g * kDragSeg, hi = min(ne, lo + kDragSeg);
    const double A = guess[w];
    const int ea = dexp(A);
    int flags = 0;
    if (ea == 0 || ea == 0x7ff) flags |= kSegNoGuess;
    const int e = ea - 1023;
    long long pmin = 0, pmax = 0;
    bool big = false, allNegZero = true;
    TieState st;
    for (long long r = lo; r < hi; r += 64) {
        const long long t = r + lane;
        const double d = t < hi ? D[c * ne + t] : __longlong_as_double((long long)kNegZero);
        allNegZero = allNegZero && dbits(d) == kNegZero;
        bool b = false, ti = false;
        long long n = 0;
        if (!(flags & kSegNoGuess)) n = seg_step(d, e, b, ti);
        big = big || b;
        long long own;
        bool after;
        const long long p = tie_row(n, ti, false, 0, st, own, after);
        pmin = min(pmin, p);
        pmax = max(pmax, p + (after ? 1 : 0));  // the first tie may add 1 (its a_f)
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        pmin = min(pmin, __shfl_xor(pmin, d, 64));
        pmax = max(pmax, __shfl_xor(pmax, d, 64));
    }
    if (__any(big)) flags |= kSegBig;
    if (__all(allNegZero)) flags |= kSegAllNegZero;
    if (st.seen) flags |= kSegHasTie;
    if (lane == 0) {
        DragSeg sg;
        sg.total = st.Q + st.A;
        sg.pmin = pmin;
        sg.pmax = pmax;
        sg.qf = st.firstQ;
        sg.e = e;
        sg.flags = flags;
        seg[w] = sg;
    }
}

// Step S: every row of kDragRow elements added element by element from a
// guessed start value, all rows at once (one thread per row): the chain's
// start for the first row; else, with Sold given, Sold's value at the row's
// first position, or, without, the segment's approximate start guess[] plus
// the approximate sums of the rows before it.  S receives the values.  With
// Sold given, specStart/specEnd receive each segment's start and end and
// specOk whether its rows chain (every row ends on the next row's guessed
// start bit for bit): then the segment is exact whenever its start is, which
// step D checks.  Without Sold this is the approximate scan of the first
// fixed-point iterations.
__global__ __launch_bounds__(256) void k_scan_spec(const double* __restrict__ D, long long ne, int chains, int nseg,
                                                   const double* __restrict__ start, const double* __restrict__ guess,
                                                   const double* __restrict__ rowSig, const double* __restrict__ Sold,
                                                   double* __restrict__ S, double* __restrict__ specStart,
                                                   double* __restrict__ specEnd, int* __restrict__ specOk) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const long long rowsPerChain = (long long)nseg * kDragRows;
    bool ok = true;
    if (t < chains * rowsPerChain) {
        const int c = (int)(t / rowsPerChain);
        const long long rem = t % rowsPerChain;
        const int g = (int)(rem / kDragRows), r = (int)(rem % kDragRows);
        const long long w = (long long)c * nseg + g;
        const long long segLo = (long long)g * kDragSeg, segHi = min(ne, segLo + kDragSeg);
        const long long lo = segLo + (long long)r * kDragRow, hi = min(segHi, lo + kDragRow);
        const int lastRow = (int)((segHi - segLo - 1) / kDragRow);
        if (r <= lastRow) {
            double v;
            if (g == 0 && r == 0) {
                v = start[c];
            } else if (Sold) {
                v = Sold[c * ne + lo];
            } else {
                v = guess[w];
                for (int q = 0; q < r; q++) v += rowSig[w * kDragRows + q];
            }
            if (Sold && r == 0) specStart[w] = v;
            const double* __restrict__ d = D + c * ne;
            double* __restrict__ out = S + c * ne;
            constexpr int kB = 16;  // loads issued ahead of the dependent adds
            long long e0 = lo;
            for (; e0 + kB <= hi; e0 += kB) {
                double x[kB];
#pragma unroll
                for (int k = 0; k < kB; k++) x[k] = d[e0 + k];
#pragma unroll
                for (int k = 0; k < kB; k++) {
                    out[e0 + k] = v;
                    v = v + x[k];
                }
            }
            for (long long e = e0; e < hi; e++) {
                out[e] = v;
                v = v + d[e];
            }
            if (Sold) {
                if (r < lastRow)
                    ok = dbits(v) == dbits(Sold[c * ne + hi]);
                else
                    specEnd[w] = v;
            }
        }
    }
    // the rows of a segment are 16 consecutive lanes, aligned to 16
    const unsigned long long bad = __ballot(!ok);
    if (Sold && (lane & (kDragRows - 1)) == 0 && t < chains * rowsPerChain)
        specOk[t / kDragRows] = ((bad >> (lane & ~(kDragRows - 1))) & ((1ULL << kDragRows) - 1)) == 0;
}

// Step D: the exact chain over the segments, one workgroup per chain.
// Thread 0 carries the exact value.  A segment is taken from step S when its
// rows chain and its guessed start equals that value bit for bit, else from
// its integer summary
// when that is valid from the value, else it is staged in LDS and done by
// wavefront 0 in blocks of 64 (chain_block: most blocks stay in one binade
// even where the segment does not), writing the values the bodies see when S
// is given.  segStart/segMode receive each segment's exact start value and
// how it was done (0 integer, 1 unchanged, 2 in blocks, 3 from step S);
// end[c] the chain's final value; counters[0] += segments done in blocks.
__global__ __launch_bounds__(256) void k_scan_chain(const double* __restrict__ D, long long ne, int nseg,
                                                    const DragSeg* __restrict__ seg, const double* __restrict__ start,
                                                    const double* __restrict__ specStart,
                                                    const double* __restrict__ specEnd, const int* __restrict__ specOk,
                                                    double* __restrict__ S,
                                                    double* __restrict__ segStart, int* __restrict__ segMode,
                                                    double* __restrict__ end, unsigned long long* __restrict__ counters) {
    __shared__ DragSeg tile[256];
    __shared__ double tss[256], tse[256];
    __shared__ int tok[256];
    __shared__ double buf[kDragSeg], sbuf[kDragSeg];
    __shared__ int slowAt;
    __shared__ double carried;
    const int c = blockIdx.x, tid = threadIdx.x;
    const long long two52 = 1LL << 52, two53 = 1LL << 53, two54 = 1LL << 54;
    double s = start[c];
    unsigned long long slowCount = 0;
    for (int base = 0; base < nseg; base += 256) {
        const int cnt = min(256, nseg - base);
        if (tid < cnt) {
            const long long gi = (long long)c * nseg + base + tid;
            tile[tid] = seg[gi];
            tss[tid] = specStart ? specStart[gi] : 0.0;
            tse[tid] = specStart ? specEnd[gi] : 0.0;
            tok[tid] = specStart ? specOk[gi] : 0;
        }
        __syncthreads();
        int pos = 0;
        while (pos < cnt) {
            if (tid == 0) {
                int q = pos;
                for (; q < cnt; q++) {
                    const DragSeg ss = tile[q];
                    const long long gi = (long long)c * nseg + base + q;
                    segStart[gi] = s;
                    if (tok[q] && dbits(tss[q]) == dbits(s)) {
                        segMode[gi] = 3;
                        s = tse[q];
                        continue;
                    }
                    if (ss.flags & kSegAllNegZero) {
                        segMode[gi] = 1;
                        continue;
                    }
                    const int es = dexp(s);
                    bool ok = !(ss.flags & (kSegNoGuess | kSegBig)) && es != 0 && es != 0x7ff &&
                              es - 1023 == ss.e && ss.pmin >= -two54 && ss.pmax <= two54;
                    if (ok) {
                        const long long M = (long long)ldexp(s, 52 - ss.e);  // exact: |M| in [2^52, 2^53)
                        const long long lo = M + ss.pmin, hi = M + ss.pmax;  // bounds of every partial value
                        ok = M > 0 ? (lo > two52 && hi < two53) : (lo > -two53 && hi < -two52);
                        if (ok) {
                            const long long af = (ss.flags & kSegHasTie) ? ((M + ss.qf) & 1) : 0;
                            segMode[gi] = 0;
                            s = ldexp((double)(M + ss.total + af), ss.e - 52);
                        }
                    }
                    if (!ok) {
                        segMode[gi] = 2;
                        break;
                    }
                }
                slowAt = q;
                carried = s;
            }
            __syncthreads();
            const int q = slowAt;
            if (q < cnt) {
                const long long lo = (long long)(base + q) * kDragSeg;
                const int len = (int)min((long long)kDragSeg, ne - lo);
                for (int t = tid; t < len; t += blockDim.x) buf[t] = D[c * ne + lo + t];
                __syncthreads();
                if (tid < 64) {
                    double v = carried;
                    for (int b0 = 0; b0 < len; b0 += 64) {
                        const int bl = min(64, len - b0);
                        const double d = tid < bl ? buf[b0 + tid] : __longlong_as_double((long long)kNegZero);
                        v = chain_block(v, d, bl, S ? sbuf + b0 : nullptr);
                    }
                    if (tid == 0) {
                        s = v;
                        slowCount++;
                    }
                }
                __syncthreads();
                if (S)
                    for (int t = tid; t < len; t += blockDim.x) S[c * ne + lo + t] = sbuf[t];
                pos = q + 1;
            } else {
                pos = cnt;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        end[c] = s;
        if (slowCount) atomicAdd(counters, slowCount);
    }
}

// Step E: the values the bodies see, S[c][t] (exclusive prefix), for the
// integer and unchanged segments (the others were written by steps S and D),
// and the comparison with the previous guess Sold (null: the start values):
// *changed = 1 if any differs bit for bit.
__global__ __launch_bounds__(256) void k_scan_fill(const double* __restrict__ D, long long ne, int chains, int nseg,
                                                   const DragSeg* __restrict__ seg,
                                                   const double* __restrict__ segStart,
                                                   const int* __restrict__ segMode, const double* __restrict__ start,
                                                   const double* __restrict__ Sold, double* __restrict__ S,
                                                   int* __restrict__ changed) {
    const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= (long long)chains * nseg) return;
    const int c = (int)(w / nseg), g = (int)(w % nseg);
    const long long lo = (long long)g * kDragSeg, hi = min(ne, lo + kDragSeg);
    const int mode = segMode[w];
    const double s0 = segStart[w];
    const DragSeg ss = seg[w];
    const long long M = mode == 0 ? (long long)ldexp(s0, 52 - ss.e) : 0;
    TieState st;
    bool diff = false;
    for (long long r = lo; r < hi; r += 64) {
        const long long t = r + lane;
        double v;
        if (mode == 0) {
            const double d = t < hi ? D[c * ne + t] : 0.0;
            bool b = false, ti = false;
            const long long n = t < hi ? seg_step(d, ss.e, b, ti) : 0;
            long long own;
            bool after;
            const long long p = tie_row(n, ti && t < hi, true, M, st, own, after);
            v = ldexp((double)(M + p - own), ss.e - 52);
        } else if (mode == 1) {
            v = s0;
        } else {
            v = t < hi ? S[c * ne + t] : 0.0;
        }
        if (t < hi) {
            if (mode < 2) S[c * ne + t] = v;
            const double o = Sold ? Sold[c * ne + t] : start[c];
            diff = diff || dbits(o) != dbits(v);
        }
    }
    if (__any(diff) && lane == 0) atomicOr(changed, 1);
}

}  // namespace gx
