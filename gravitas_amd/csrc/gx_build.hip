// gx_build.hip -- tree construction on gfx950.
//
// Compiled with -ffp-contract=off: every FP64 expression here is evaluated
// with the same operations in the same order as the reference, so keys,
// permutation, node set and moments are bit-identical to it.
//
//   collect          internal/sim/bh/build.go:62-113
//   keys             internal/sim/bh/build.go:19-31, morton.go:7-21
//   radix sort       internal/sim/bh/morton.go:25-49 (stable LSD, 8-bit digits)
//   sorted copies    internal/sim/bh/build.go:34-41, internal/sim/sim.go:93-98
//   topology         internal/sim/bh/build.go:134-172, lowerBound :340-350
//   moments          internal/sim/bh/build.go:176-217, quadrupole.add :326-337
//
// The reference builds top-down recursively; here the node set follows from
// the sorted keys in four launches at any depth (each star's chain of nodes
// from the keys within leafSize+1 of it, see the topology section), and
// moments are one launch for all leaves, then one per level from the deepest
// level up.  Children of a node are stored contiguously and in octant order,
// and skip links follow the reference's rule, so the walk order -- which
// alone determines the forces -- equals the reference's.  Only the numbering
// of nodes differs (level by level instead of depth-first).
#include <hip/hip_fp16.h>

#include "gx_common.h"
#include "gx_kernels.h"

namespace gx {

// ---------------------------------------------------------------------------
// collect
// ---------------------------------------------------------------------------
// Zeroes the step's counters and the per-instance accumulators.
__global__ void k_step_init(Meta* meta, Batch batch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) {
        for (int l = 0; l <= kMaxLevels; l++) {
            meta->levelStart[l] = 0;
            meta->levelCount[l] = 0;
        }
        meta->nstars = meta->nholes = meta->nnodes = 0;
        meta->pp = meta->pc = 0;
        meta->flags = 0;
        meta->maxStack = 0;
        meta->maxList[0] = meta->maxList[1] = meta->maxList[2] = 0;
        meta->listEntries = meta->rounds = meta->mixed = 0;
        for (int q = 0; q < 8; q++) meta->xcdNext[q] = 0;
        meta->ntargets = 0;
        meta->topoWhy = 0;
        meta->topoNode = 1ll << 62;
    }
    if (b < batch.B) {
        for (int c = 0; c < 3; c++) {
            batch.boxEnc[6 * b + c] = enc_d(INFINITY);
            batch.boxEnc[6 * b + 3 + c] = enc_d(-INFINITY);
        }
        batch.starCount[b] = 0;
    }
}

__device__ __forceinline__ void inst_accumulate(const Batch& batch, int b, int cnt, const double mn[3],
                                                const double mx[3]) {
    if (cnt == 0) return;
    atomicAdd(&batch.starCount[b], (unsigned long long)cnt);
    for (int c = 0; c < 3; c++) {
        atomicMin(&batch.boxEnc[6 * b + c], enc_d(mn[c]));
        atomicMax(&batch.boxEnc[6 * b + 3 + c], enc_d(mx[c]));
    }
}

__global__ __launch_bounds__(256) void k_collect_count(const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       const uint8_t* __restrict__ kind,
                                                       const uint8_t* __restrict__ alive, long long n,
                                                       int* __restrict__ blockStar, int* __restrict__ blockHole,
                                                       double* __restrict__ blockBox, Batch batch) {
    __shared__ double sbox[6][4];
    __shared__ int scnt[2][4];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    long long base = (long long)blockIdx.x * kCollectTile;
    int cnt = 0, hcnt = 0;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int j = 0; j < kCollectTile / 256; j++) {
        long long i = base + j * 256 + tid;
        if (i >= n) break;
        bool a = alive[i] != 0;
        uint8_t k = kind[i];
        if (a && k == kHole) hcnt++;
        if (a && k != kHole) {
            cnt++;
            double px = x[i], py = y[i], pz = z[i];
            mn[0] = fmin(mn[0], px); mx[0] = fmax(mx[0], px);
            mn[1] = fmin(mn[1], py); mx[1] = fmax(mx[1], py);
            mn[2] = fmin(mn[2], pz); mx[2] = fmax(mx[2], pz);
        }
    }
    // instances of the tile's first and last body: one instance -> block-wide
    // reduction and one set of atomics; several -> each wave reduces every
    // instance of each of its 64-body chunks and flushes one set of atomics per
    // instance and chunk (per-star atomics on a few shared addresses serialise)
    const long long last = min(n, base + kCollectTile) - 1;
    const int b0 = last >= base ? upper_index(batch.bodyOff, batch.B, base) : 0;
    const int b1 = last >= base ? upper_index(batch.bodyOff, batch.B, last) : 0;
    if (b0 != b1) {
        for (int j = 0; j < kCollectTile / 256; j++) {
            const long long c0 = base + j * 256 + w * 64;  // the wave's chunk
            if (c0 >= n) break;
            const long long c1 = min(n, c0 + 64) - 1, i = c0 + lane;
            const bool in = i <= c1 && alive[i] && kind[i] != kHole;
            double p[3] = {0.0, 0.0, 0.0};
            int bi = -1;
            if (in) {
                p[0] = x[i];
                p[1] = y[i];
                p[2] = z[i];
                bi = upper_index(batch.bodyOff, batch.B, i);
            }
            const int lo = upper_index(batch.bodyOff, batch.B, c0), hi = upper_index(batch.bodyOff, batch.B, c1);
            for (int b = lo; b <= hi; b++) {
                const bool m = bi == b;
                const int cc = wave_sum_i(m ? 1 : 0);
                if (cc == 0) continue;
                double bmn[3], bmx[3];
                for (int c = 0; c < 3; c++) {
                    bmn[c] = wave_min_d(m ? p[c] : INFINITY);
                    bmx[c] = wave_max_d(m ? p[c] : -INFINITY);
                }
                if (lane == 0) inst_accumulate(batch, b, cc, bmn, bmx);
            }
        }
    }
    cnt = wave_sum_i(cnt);
    hcnt = wave_sum_i(hcnt);
    for (int c = 0; c < 3; c++) {
        mn[c] = wave_min_d(mn[c]);
        mx[c] = wave_max_d(mx[c]);
    }
    if (lane == 0) {
        scnt[0][w] = cnt;
        scnt[1][w] = hcnt;
        for (int c = 0; c < 3; c++) {
            sbox[c][w] = mn[c];
            sbox[3 + c][w] = mx[c];
        }
    }
    __syncthreads();
    if (tid == 0) {
        int c0 = 0, c1 = 0;
        double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int ww = 0; ww < 4; ww++) {
            c0 += scnt[0][ww];
            c1 += scnt[1][ww];
            for (int c = 0; c < 3; c++) {
                b[c] = fmin(b[c], sbox[c][ww]);
                b[3 + c] = fmax(b[3 + c], sbox[3 + c][ww]);
            }
        }
        blockStar[blockIdx.x] = c0;
        blockHole[blockIdx.x] = c1;
        if (batch.B == 1) {  // one instance: k_collect_scan reduces the blocks' boxes (no contended atomics)
            for (int c = 0; c < 6; c++) blockBox[6 * (size_t)blockIdx.x + c] = b[c];
        } else if (b0 == b1) {
            inst_accumulate(batch, b0, c0, b, b + 3);
        }
    }
}

__device__ __forceinline__ double box_side(const Batch& batch, int b) {
    const double mnx = dec_d(batch.boxEnc[6 * b]), mny = dec_d(batch.boxEnc[6 * b + 1]);
    const double mnz = dec_d(batch.boxEnc[6 * b + 2]), mxx = dec_d(batch.boxEnc[6 * b + 3]);
    const double mxy = dec_d(batch.boxEnc[6 * b + 4]), mxz = dec_d(batch.boxEnc[6 * b + 5]);
    double side = fmax(mxx - mnx, fmax(mxy - mny, mxz - mnz));
    return fmax(side * (1 + 1e-9), 1e-12);  // build.go:19-20
}

// Exclusive scan of per-thread values v (1024 threads) in LDS; returns the
// prefix before this thread and the total in *total.
// Exclusive scans of K values per thread over a 1,024-thread workgroup, all
// K at once: shuffles within each wavefront, then the 16 wavefront totals
// (two barriers instead of twenty per value).  v becomes the exclusive
// prefix, total the sum over the workgroup.
template <int K>
__device__ void block_scan_1024(long long (&v)[K], long long (&total)[K],
                                long long (*ws)[kScanThreads / kWave + 1]) {
    constexpr int W = kScanThreads / kWave;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    long long incl[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        incl[k] = v[k];
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const long long t = __shfl_up(incl[k], d, kWave);
            if (lane >= d) incl[k] += t;
        }
        if (lane == kWave - 1) ws[k][w] = incl[k];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const long long x = lane < W ? ws[k][lane] : 0;
            long long xi = x;
#pragma unroll
            for (int d = 1; d < W; d <<= 1) {
                const long long t = __shfl_up(xi, d, kWave);
                if (lane >= d) xi += t;
            }
            if (lane < W) ws[k][lane] = xi - x;  // exclusive prefix of the wavefront totals
            if (lane == W - 1) ws[k][W] = xi;    // the workgroup total
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        total[k] = ws[k][W];
        v[k] = ws[k][w] + incl[k] - v[k];
    }
    __syncthreads();  // ws is reused by the next scan
}

// Single workgroup: exclusive scans of the per-block counts (the stable
// compaction of collect) and of the per-instance star and group counts;
// instance roots and root cells.
__global__ __launch_bounds__(1024) void k_collect_scan(int nb, int* __restrict__ blockStar,
                                                       int* __restrict__ blockHole, const double* __restrict__ blockBox,
                                                       Meta* meta, Batch batch) {
    constexpr int W = kScanThreads / kWave;
    __shared__ long long ws[3][W + 1];
    __shared__ double sb[6][W];
    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    {
        // wave w takes the blocks [w*chunk, (w+1)*chunk) in slices of 64
        // consecutive ones (coalesced loads, a wavefront scan per slice)
        const int chunk = (nb + W - 1) / W;
        const int lo = min(nb, w * chunk), hi = min(nb, lo + chunk);
        long long v[2] = {0, 0}, total[2];
        double bx[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int i0 = lo; i0 < hi; i0 += kWave) {
            const int i = i0 + lane;
            if (i < hi) {
                v[0] += blockStar[i];
                v[1] += blockHole[i];
                if (batch.B == 1)
                    for (int c = 0; c < 3; c++) {
                        bx[c] = fmin(bx[c], blockBox[6 * (size_t)i + c]);
                        bx[3 + c] = fmax(bx[3 + c], blockBox[6 * (size_t)i + 3 + c]);
                    }
            }
        }
        if (batch.B == 1) {  // the one instance's box: wave, then workgroup reduction
            for (int c = 0; c < 3; c++) {
                bx[c] = wave_min_d(bx[c]);
                bx[3 + c] = wave_max_d(bx[3 + c]);
            }
            if (lane == 0)
                for (int c = 0; c < 6; c++) sb[c][w] = bx[c];
        }
        // the waves' totals, scanned across the workgroup (lane sums of a
        // wave reduce to its total; the wave's own order is rebuilt below)
        long long t[2] = {(long long)wave_sum_u64((unsigned long long)v[0]),
                          (long long)wave_sum_u64((unsigned long long)v[1])};
        if (lane == 0) {
            ws[0][w] = t[0];
            ws[1][w] = t[1];
        }
        __syncthreads();
        if (w == 0) {  // the first wave scans the wave totals and reduces the wave boxes
            const long long c0 = lane < W ? ws[0][lane] : 0, c1 = lane < W ? ws[1][lane] : 0;
            long long p0 = c0, p1 = c1;
#pragma unroll
            for (int d = 1; d < W; d <<= 1) {
                const long long t0 = __shfl_up(p0, d, kWave), t1 = __shfl_up(p1, d, kWave);
                if (lane >= d) {
                    p0 += t0;
                    p1 += t1;
                }
            }
            if (lane < W) {
                ws[0][lane] = p0 - c0;
                ws[1][lane] = p1 - c1;
            }
            if (lane == W - 1) {
                ws[0][W] = p0;
                ws[1][W] = p1;
            }
            if (batch.B == 1)
                for (int c = 0; c < 6; c++) {
                    const double v = lane < W ? sb[c][lane] : (c < 3 ? INFINITY : -INFINITY);
                    const double r = c < 3 ? wave_min_d(v) : wave_max_d(v);
                    if (lane == 0) batch.boxEnc[c] = enc_d(r);
                }
        }
        __syncthreads();
        total[0] = ws[0][W];
        total[1] = ws[1][W];
        if (tid == 0) {
            meta->nstars = meta->ntargets = total[0];
            meta->nholes = total[1];
            if (batch.B == 1) batch.starCount[0] = (unsigned long long)total[0];
        }
        // exclusive offsets, slice by slice in block order
        long long run0 = ws[0][w], run1 = ws[1][w];
        for (int i0 = lo; i0 < hi; i0 += kWave) {
            const int i = i0 + lane;
            const int c0 = i < hi ? blockStar[i] : 0, c1 = i < hi ? blockHole[i] : 0;
            const int p0 = wave_incl_scan(c0), p1 = wave_incl_scan(c1);
            if (i < hi) {
                blockStar[i] = (int)(run0 + p0 - c0);
                blockHole[i] = (int)(run1 + p1 - c1);
            }
            run0 += __shfl(p0, kWave - 1, kWave);
            run1 += __shfl(p1, kWave - 1, kWave);
        }
    }
    __syncthreads();  // starCount[0] (one instance) is read below
    {
        const int B = batch.B;
        const int per = (B + 1023) / 1024;
        const int lo = tid * per, hi = min(B, lo + per);
        long long v[3] = {0, 0, 0}, total[3];  // stars, groups, roots
        for (int b = lo; b < hi; b++) {
            const long long c = (long long)batch.starCount[b];
            v[0] += c;
            v[1] += (c + batch.gsize - 1) / batch.gsize;
            v[2] += c > 0;
        }
        block_scan_1024<3>(v, total, ws);
        if (tid == 0) {
            batch.starOff[B] = total[0];
            batch.groupOff[B] = total[1];
            meta->levelCount[0] = total[2];
        }
        long long rs = v[0], rg = v[1], rr = v[2];
        for (int b = lo; b < hi; b++) {
            const long long c = (long long)batch.starCount[b];
            batch.starOff[b] = rs;
            batch.groupOff[b] = rg;
            batch.root[b] = c > 0 ? (int)rr : -1;
            batch.half[b] = c > 0 ? box_side(batch, b) / 2 : 0.0;
            rs += c;
            rg += (c + batch.gsize - 1) / batch.gsize;
            rr += c > 0;
        }
    }
}

// ---------------------------------------------------------------------------
// keys: build.go:19-31 (written by k_collect_scatter at each star's slot)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t cell_of(double v, double lo, double scale) {
    double t = fmax(0.0, (v - lo) * scale);
    return (uint64_t)fmin(2097151.0, t);
}

__device__ __forceinline__ uint64_t morton_key(double x, double y, double z, const Batch& batch, int b) {
    const double side = box_side(batch, b);
    const double scale = (double)(1 << kMortonBits) / side;
    const uint64_t cx = cell_of(x, dec_d(batch.boxEnc[6 * b]), scale);
    const uint64_t cy = cell_of(y, dec_d(batch.boxEnc[6 * b + 1]), scale);
    const uint64_t cz = cell_of(z, dec_d(batch.boxEnc[6 * b + 2]), scale);
    return spread21(cx) << 2 | spread21(cy) << 1 | spread21(cz);
}

// Stable compaction in state order: order[] gets the stars, holes[] the
// holes, and keys[] each star's Morton key at its slot (the boxes are final).
// Body i = base + j*256 + tid (coalesced); ranks from ballots, in body order.
__global__ __launch_bounds__(256) void k_collect_scatter(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         const uint8_t* __restrict__ kind,
                                                         const uint8_t* __restrict__ alive, long long n,
                                                         const int* __restrict__ starOff,
                                                         const int* __restrict__ holeOff, Batch batch,
                                                         int32_t* __restrict__ order, int32_t* __restrict__ holes,
                                                         uint64_t* __restrict__ keys, uint16_t* __restrict__ instOf) {
    constexpr int per = kCollectTile / 256, W = 256 / kWave;
    __shared__ int wcnt[2][per][W];
    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    const long long base = (long long)blockIdx.x * kCollectTile;
    uint64_t bs[per], bh[per];
#pragma unroll
    for (int j = 0; j < per; j++) {
        const long long i = base + j * 256 + tid;
        const bool a = i < n && alive[i];
        const bool hole = a && kind[i] == kHole;
        bs[j] = __ballot(a && !hole);
        bh[j] = __ballot(hole);
        if (lane == 0) {
            wcnt[0][j][w] = __popcll(bs[j]);
            wcnt[1][j][w] = __popcll(bh[j]);
        }
    }
    __syncthreads();
    if (tid < 2) {  // exclusive prefix over (slice, wave), in body order
        int acc = tid == 0 ? starOff[blockIdx.x] : holeOff[blockIdx.x];
        for (int j = 0; j < per; j++)
            for (int ww = 0; ww < W; ww++) {
                const int c = wcnt[tid][j][ww];
                wcnt[tid][j][ww] = acc;
                acc += c;
            }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1;
#pragma unroll
    for (int j = 0; j < per; j++) {
        const long long i = base + j * 256 + tid;
        if ((bs[j] >> lane) & 1) {
            const int os = wcnt[0][j][w] + __popcll(bs[j] & below);
            const int b = batch.B > 1 ? upper_index(batch.bodyOff, batch.B, i) : 0;
            keys[os] = morton_key(x[i], y[i], z[i], batch, b);
            order[os] = (int32_t)i;
            if (batch.B > 1) instOf[i] = (uint16_t)b;
        }
        if ((bh[j] >> lane) & 1) holes[wcnt[1][j][w] + __popcll(bh[j] & below)] = (int32_t)i;
    }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, order) pairs, 8-bit digits.
// ---------------------------------------------------------------------------
// Digit of a sort element: 8 bits of the Morton key, or (instance passes, the
// last ones) 8 bits of the instance of the element's body, which the collect
// wrote per body.
__device__ __forceinline__ uint32_t sort_digit(uint64_t key, int32_t val, int shift, bool instPass,
                                               const Batch& batch) {
    if (instPass) return ((uint32_t)batch.instOf[val] >> shift) & 255u;
    return (uint32_t)(key >> shift) & 255u;
}

__global__ __launch_bounds__(256) void k_sort_hist(const uint64_t* __restrict__ keys, const int32_t* __restrict__ vals,
                                                   const Meta* __restrict__ meta, int shift, int instPass,
                                                   Batch batch, uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    long long nst = meta->nstars;
    long long base = (long long)blockIdx.x * kSortTile;
    constexpr int R = kSortTile / 256;
    uint32_t dig[R];  // all loads of the tile in flight before the first atomic
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long long i = base + r * 256 + tid;
        dig[r] = i < nst ? sort_digit(keys[i], instPass ? vals[i] : 0, shift, instPass, batch) : 256u;
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        if (dig[r] < 256u) atomicAdd(&h[dig[r]], 1u);
    __syncthreads();
    hist[(size_t)tid * ntiles + blockIdx.x] = h[tid];
}

// One workgroup per digit: exclusive scan of the digit's counts over the
// tiles; the digit's total.  Rows of up to 1,024 tiles: thread t scans the
// run [t per, (t+1) per) (per <= 4) in registers, and the run sums are scanned
// with shuffles and one LDS step.  Longer rows go in chunks of 2,048 counts:
// a chunk is loaded coalesced into LDS (padded by one word in 32 against
// bank conflicts), thread t takes the counts [8t, 8t+8), the thread sums are
// scanned as above, the offsets go back through LDS and out coalesced; a
// running carry joins the chunks.
__global__ __launch_bounds__(kRowScanThreads) void k_sort_scan(uint32_t* __restrict__ hist, int ntiles,
                                                               uint32_t* __restrict__ digitTotal) {
    constexpr int T = kRowScanThreads, K = 8, C = T * K, W = T / kWave;
    __shared__ uint32_t buf[C + C / 32];
    __shared__ uint32_t ws[W];
    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    if (ntiles <= 4 * T) {
        const int per = (ntiles + T - 1) / T;
        const int lo = min(tid * per, ntiles), hi = min(lo + per, ntiles);
        uint32_t sum = 0;
        for (int i = lo; i < hi; i++) sum += row[i];
        uint32_t incl = sum;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += t;
        }
        if (lane == kWave - 1) ws[w] = incl;
        __syncthreads();
        uint32_t off = incl - sum, total = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {
            const uint32_t x = ws[k];
            off += k < w ? x : 0u;
            total += x;
        }
        for (int i = lo; i < hi; i++) {
            const uint32_t c = row[i];
            row[i] = off;
            off += c;
        }
        if (tid == 0) digitTotal[blockIdx.x] = total;
        return;
    }
    uint32_t carry = 0;
    for (int base = 0; base < ntiles; base += C) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = k * T + tid, i = base + j;
            buf[j + (j >> 5)] = i < ntiles ? row[i] : 0u;
        }
        __syncthreads();
        uint32_t c[K], sum = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = tid * K + k;
            c[k] = buf[j + (j >> 5)];
            sum += c[k];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += t;
        }
        if (lane == kWave - 1) ws[w] = incl;
        __syncthreads();
        uint32_t off = carry + incl - sum, tot = 0;
#pragma unroll
        for (int k = 0; k < W; k++) {
            const uint32_t x = ws[k];
            off += k < w ? x : 0u;
            tot += x;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = tid * K + k;
            buf[j + (j >> 5)] = off;
            off += c[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = k * T + tid, i = base + j;
            if (i < ntiles) row[i] = buf[j + (j >> 5)];
        }
        carry += tot;
        __syncthreads();  // buf and ws are rewritten by the next chunk
    }
    if (tid == 0) digitTotal[blockIdx.x] = carry;
}

// Scatter of one tile: the tile's elements are first placed in LDS in their
// output order (by digit, stable within a digit), then written out so that
// consecutive threads write consecutive positions of each digit's run (the
// stores coalesce instead of scattering over 256 runs).  With rankOut (the
// last pass) each body's sorted position is recorded by its state index.
__global__ __launch_bounds__(256) void k_sort_scatter(const uint64_t* __restrict__ kin,
                                                      const int32_t* __restrict__ vin, uint64_t* __restrict__ kout,
                                                      int32_t* __restrict__ vout, const Meta* __restrict__ meta,
                                                      int shift, int instPass, Batch batch,
                                                      const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ digitTotal, int ntiles,
                                                      int32_t* __restrict__ rankOut) {
    constexpr int R = kSortTile / 256, W = 256 / kWave;
    __shared__ uint32_t globalBase[256];  // output index of digit d's first element of this tile
    __shared__ uint32_t localStart[256];  // staging slot of digit d's first element
    __shared__ uint16_t cnt[R][W][256];   // per (round, wave): elements of digit d, then their first slot
    __shared__ uint32_t wsum[2][W];       // per wave: sums of its 64 digits' totals and tile counts
    __shared__ uint64_t skey[kSortTile];
    __shared__ int32_t sval[kSortTile];
    __shared__ uint8_t sdig[kSortTile];
    const int tid = threadIdx.x, w = tid >> 6;
    const long long nst = meta->nstars;
    // digit bases (exclusive scan of the digit totals) and this tile's digit
    // counts (differences of the per-tile prefixes k_sort_scan left in hist)
    const uint32_t tot = digitTotal[tid];
    const uint32_t pre = hist[(size_t)tid * ntiles + blockIdx.x];
    const uint32_t tc = ((int)blockIdx.x + 1 < ntiles ? hist[(size_t)tid * ntiles + blockIdx.x + 1] : tot) - pre;
#pragma unroll
    for (int r = 0; r < R; r++)
        for (int ww = 0; ww < W; ww++) cnt[r][ww][tid] = 0;
    const long long base = (long long)blockIdx.x * kSortTile;
    uint64_t keyr[R];  // the whole tile is loaded up front: one memory latency, not R
    int32_t valr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long long idx = base + r * 256 + tid;
        keyr[r] = idx < nst ? kin[idx] : 0;
        valr[r] = idx < nst ? vin[idx] : 0;
    }
    // exclusive scans over the digits (thread = digit): shuffles within each
    // wavefront, then the wavefronts' sums (the barrier also orders the
    // zeroed counts before the ranks below)
    const int lane = lane_id();
    uint32_t a0 = tot, a1 = tc;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t0 = __shfl_up(a0, d, kWave), t1 = __shfl_up(a1, d, kWave);
        if (lane >= d) {
            a0 += t0;
            a1 += t1;
        }
    }
    if (lane == kWave - 1) {
        wsum[0][w] = a0;
        wsum[1][w] = a1;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < W; k++)
        if (k < w) {
            a0 += wsum[0][k];
            a1 += wsum[1][k];
        }
    globalBase[tid] = a0 - tot + pre;
    localStart[tid] = a1 - tc;
    // ranks within each (round, wave) from ballots: all rounds first, so one
    // barrier, one pass over the (round, wave) counts per digit and one more
    // barrier place the whole tile (stable: round, wave, lane order)
    uint32_t digr[R];
    int rankr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long long idx = base + r * 256 + tid;
        const bool valid = idx < nst;
        const uint32_t dig = valid ? sort_digit(keyr[r], valr[r], shift, instPass, batch) : 0u;
        uint64_t match = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            match &= bit ? bal : ~bal;
        }
        rankr[r] = valid ? mbcnt(match) : -1;
        digr[r] = dig;
        if (valid && rankr[r] == 0) cnt[r][w][dig] = (uint16_t)__popcll(match);
    }
    __syncthreads();
    {
        uint32_t run = localStart[tid];
        for (int r = 0; r < R; r++)
            for (int ww = 0; ww < W; ww++) {
                const uint32_t c = cnt[r][ww][tid];
                cnt[r][ww][tid] = (uint16_t)run;
                run += c;
            }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++)
        if (rankr[r] >= 0) {
            const uint32_t slot = (uint32_t)cnt[r][w][digr[r]] + (uint32_t)rankr[r];
            skey[slot] = keyr[r];
            sval[slot] = valr[r];
            sdig[slot] = (uint8_t)digr[r];
        }
    __syncthreads();
    const int count = (int)min((long long)kSortTile, nst - base);
    for (int j = tid; j < count; j += 256) {
        const uint32_t d = sdig[j];
        const uint32_t dst = globalBase[d] + ((uint32_t)j - localStart[d]);
        const int32_t val = sval[j];
        kout[dst] = skey[j];
        vout[dst] = val;
        if (rankOut) rankOut[val] = (int32_t)dst;  // last pass: state index -> sorted position
    }
}

// ---------------------------------------------------------------------------
// sorted copies: build.go:34-41 with SourceMass (sim.go:93-98)
// ---------------------------------------------------------------------------
// The sorted copies are written from the state side: body i (read in state
// order, coalesced) goes to its sorted position rank[i], written by the last
// sort pass.  Reading the state through the sorted order instead fetches a
// whole cache line for every 8-byte read (7x the bytes at 16M bodies).
__global__ __launch_bounds__(256) void k_scatter_spos(const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, const double* __restrict__ m,
                                                      const uint8_t* __restrict__ kind,
                                                      const uint8_t* __restrict__ alive, long long n,
                                                      const int32_t* __restrict__ rank, int dmOff,
                                                      double4* __restrict__ spos) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint8_t k = kind[i];
        if (!alive[i] || k == kHole) continue;  // the stars of collect (build.go:62-113)
        const double sm = (dmOff && k == kDark) ? 0.0 : m[i];
        spos[rank[i]] = make_double4(x[i], y[i], z[i], sm);
    }
}

// ---------------------------------------------------------------------------
// topology: buildNode/split, build.go:134-172 (lowerBound :340-350)
// ---------------------------------------------------------------------------
// The reference splits a node of more than leafSize stars into its non-empty
// octants, down to level mortonBits.  Restated over the sorted keys: let
// delta_i be the number of leading 3-bit digits keys i and i+1 share (-1
// between two instances and past the ends).  The stars sharing a level-L
// prefix are a run whose interior deltas are all >= L (a "level-L run"); a
// run is a node when its parent run (level L-1) has more than leafSize stars,
// and a node splits when it has more than leafSize stars and L < mortonBits.
// A run of more than leafSize stars holding star s contains a window of
// leafSize+1 consecutive stars that holds s, so with
//   m(j) = min(delta_j .. delta_{j+leafSize-1})   (stars j .. j+leafSize share m(j) digits)
//   M(s) = max(m(j), j = s-leafSize .. s)
// the level-L run holding s has more than leafSize stars iff L <= M(s).  The
// runs starting at star s are those of levels delta_{s-1}+1 .. mortonBits, so
// the nodes starting at s are a chain over the levels
//   Llo(s) = delta_{s-1} + 1  ..  Lhi(s) = min(M(s) + 1, mortonBits),
// every one of them split except the last.  Each star's chain follows from
// the keys within leafSize+1 of it, so the whole topology takes four
// launches at any depth instead of one per level:
//   k_topo_count   chains (Llo, Lhi) of a tile of stars; per-level node counts
//   k_sort_scan    per level, the exclusive scan of the counts over the tiles
//   k_topo_nodes   node records: level by level, each level in star order
//                  (children of a node are contiguous and in octant order);
//                  centres by the reference's halving from the root centre;
//                  leaf ends from the deltas
//   k_topo_links   split nodes' ends (a search on the keys), skip links (the
//                  node after a subtree is the chain that starts at its end),
//                  child counts and parents.
constexpr int kTopoPre = kLeafSize;            // deltas before the tile: delta_{ts-24} ..
constexpr int kTopoDeltas = kTopoTile + kTopoPre + kLeafSize - 1;  // .. delta_{te+22}
constexpr int kTopoMins = kTopoTile + kTopoPre;                    // m(j), j = ts-24 .. te-1
constexpr uint16_t kNoChain = kMaxLevels;      // Llo = 22 > Lhi = 0: no node starts here

struct TopoLDS {
    uint64_t key[kTopoDeltas + 1];  // keys ts-24 .. te+23
    int8_t delta[kTopoDeltas];
    uint8_t brk[kTopoDeltas];       // an instance starts right after delta index w
};

// The sliding windows of k_topo_count (van Herk / Gil-Werman: extremes from
// the start and from the end of fixed blocks of the window's length, so each
// window is one or two lookups).
struct TopoWin {
    int8_t pre[kTopoDeltas], suf[kTopoDeltas];  // min within blocks of leafSize deltas
    int8_t mwin[kTopoMins];                     // m(j)
    int8_t mpre[kTopoMins], msuf[kTopoMins];    // max within blocks of leafSize+1 m values
};

// Deltas of the window around tile ts (delta index w = i - (ts - 24)).
__device__ void topo_deltas(TopoLDS& S, const uint64_t* __restrict__ keys, long long n, long long ts,
                            const Batch& batch) {
    const int tid = threadIdx.x;
    const long long i0 = ts - kTopoPre;
    for (int w = tid; w <= kTopoDeltas; w += 256) {
        const long long i = i0 + w;
        S.key[w] = i >= 0 && i < n ? keys[i] : 0ull;
        if (w < kTopoDeltas) S.brk[w] = 0;
    }
    __syncthreads();
    if (batch.B > 1) {
        // instance starts i+1 = starOff[b], b = 1 .. B-1, inside the window
        const long long v = i0;  // first b with starOff[b] > i0
        int b = v < 0 ? 1 : upper_index(batch.starOff, batch.B + 1, v) + 1;
        b = max(b, 1);
        for (int bb = b + tid; bb < batch.B; bb += 256) {
            const long long w = batch.starOff[bb] - 1 - i0;
            if (w >= kTopoDeltas) break;
            if (w >= 0) S.brk[w] = 1;
        }
        __syncthreads();
    }
    for (int w = tid; w < kTopoDeltas; w += 256) {
        const long long i = i0 + w;
        int d = -1;
        if (i >= 0 && i < n - 1 && !S.brk[w]) {
            const uint64_t x = S.key[w] ^ S.key[w + 1];
            d = x ? (__clzll((long long)x) - 1) / 3 : kMortonBits;  // 63-bit keys: bit 63 is 0
        }
        S.delta[w] = (int8_t)d;
    }
    __syncthreads();
}

__device__ __forceinline__ void topo_fail(Meta* meta, int why, long long node) {
    atomicOr(&meta->flags, (int)kFlagTopology);
    atomicOr(&meta->topoWhy, why);
    atomicMin(&meta->topoNode, node);
}

// m(j) for j = ts-24 .. te-1 (window index u = j - (ts-24)), then the block
// maxima that give M(s) = max(m(s-24) .. m(s)).
__device__ void topo_windows(const TopoLDS& S, TopoWin& V) {
    const int tid = threadIdx.x;
    constexpr int BW = kLeafSize, BM = kLeafSize + 1;
    for (int b = tid; b * BW < kTopoDeltas; b += 256) {
        const int a = b * BW, e = min(a + BW, kTopoDeltas);
        int v = 127;
        for (int i = a; i < e; i++) V.pre[i] = (int8_t)(v = min(v, (int)S.delta[i]));
        v = 127;
        for (int i = e - 1; i >= a; i--) V.suf[i] = (int8_t)(v = min(v, (int)S.delta[i]));
    }
    __syncthreads();
    for (int u = tid; u < kTopoMins; u += 256)  // min(delta[u .. u+23])
        V.mwin[u] = u % BW == 0 ? V.suf[u] : (int8_t)min((int)V.suf[u], (int)V.pre[u + BW - 1]);
    __syncthreads();
    for (int b = tid; b * BM < kTopoMins; b += 256) {
        const int a = b * BM, e = min(a + BM, kTopoMins);
        int v = -128;
        for (int i = a; i < e; i++) V.mpre[i] = (int8_t)(v = max(v, (int)V.mwin[i]));
        v = -128;
        for (int i = e - 1; i >= a; i--) V.msuf[i] = (int8_t)(v = max(v, (int)V.mwin[i]));
    }
    __syncthreads();
}

// Chain of star s = ts + u (u < kTopoTile), from the window's deltas and M(s).
__device__ __forceinline__ uint16_t topo_chain(const TopoLDS& S, const TopoWin& V, int u, bool star) {
    if (!star) return kNoChain;
    constexpr int BM = kLeafSize + 1;
    const int M = u % BM == 0 ? V.msuf[u] : max((int)V.msuf[u], (int)V.mpre[u + BM - 1]);  // m(s-24) .. m(s)
    const int lo = S.delta[u + kTopoPre - 1] + 1;
    const int hi = min(M + 1, kMortonBits);
    return lo <= hi ? (uint16_t)(lo | hi << 8) : kNoChain;
}

__global__ __launch_bounds__(256) void k_topo_count(const uint64_t* __restrict__ keys, Meta* __restrict__ meta,
                                                    Batch batch, uint16_t* __restrict__ chain,
                                                    uint32_t* __restrict__ tileCount, int ntiles) {
    __shared__ TopoLDS S;
    __shared__ TopoWin V;
    __shared__ uint32_t cnt[kMaxLevels];
    const int tid = threadIdx.x;
    const long long n = meta->nstars;
    const long long ts = (long long)blockIdx.x * kTopoTile;
    if (tid < kMaxLevels) cnt[tid] = 0;
    if (ts < n) {
        topo_deltas(S, keys, n, ts, batch);
        // the runs assume keys sorted within each instance
        for (int u = tid; u < kTopoTile; u += 256) {
            const int w = u + kTopoPre;
            if (ts + u + 1 < n && !S.brk[w] && S.key[w] > S.key[w + 1]) topo_fail(meta, 64, ts + u);
        }
        topo_windows(S, V);
        for (int r = 0; r < kTopoTile / 256; r++) {
            const int u = r * 256 + tid;
            const long long s = ts + u;
            const uint16_t c = topo_chain(S, V, u, s < n);
            if (s < n) chain[s] = c;
            const int lo = c & 255, hi = c >> 8;
            for (int L = lo; L <= hi; L++) atomicAdd(&cnt[L], 1u);  // about one node per chain
        }
    }
    __syncthreads();
    if (tid < kMaxLevels) tileCount[(size_t)tid * ntiles + blockIdx.x] = cnt[tid];
}

__global__ __launch_bounds__(256) void k_topo_nodes(const uint64_t* __restrict__ keys,
                                                    const uint16_t* __restrict__ chain,
                                                    const uint32_t* __restrict__ tileOff,
                                                    const uint32_t* __restrict__ levelTotal, Meta* meta, Batch batch,
                                                    TravNode* __restrict__ trav, int32_t* __restrict__ chainFirst,
                                                    long long cap, int ntiles) {
    constexpr int R = kTopoTile / 256, W = 256 / kWave;
    __shared__ TopoLDS S;
    __shared__ uint32_t list[kTopoTile];         // the tile's chains in star order: u << 16 | chain
    __shared__ uint32_t wpos[R][W + 1];          // compaction offsets of the (r, wave) runs
    __shared__ uint32_t pre[R][W][kMaxLevels];  // exclusive prefix of the (chunk, wave) counts, in list order
    __shared__ long long base[kMaxLevels + 1];  // level start + this tile's offset in the level
    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    if (tid == 0) {
        long long acc = 0;
        for (int L = 0; L < kMaxLevels; L++) {
            base[L] = acc;
            acc += levelTotal[L];
        }
        base[kMaxLevels] = acc;
        if (blockIdx.x == 0) {
            for (int L = 0; L < kMaxLevels; L++) {
                meta->levelStart[L] = base[L];
                meta->levelCount[L] = levelTotal[L];
            }
            meta->levelStart[kMaxLevels] = acc;
            meta->levelCount[kMaxLevels] = 0;
            meta->nnodes = acc;
            if (acc > cap) atomicOr(&meta->flags, (int)kFlagNodeOverflow);
        }
    }
    __syncthreads();
    const long long n = meta->nstars;
    const long long ts = (long long)blockIdx.x * kTopoTile;
    if (base[kMaxLevels] > cap || ts >= n) return;
    if (tid < kMaxLevels) base[tid] += tileOff[(size_t)tid * ntiles + blockIdx.x];
    topo_deltas(S, keys, n, ts, batch);  // (its first barrier also publishes base)
    // the stars that start a chain (about one in seven), compacted in star
    // order, so that the level loops below run over them only
    const uint64_t below = (1ull << lane) - 1;
    uint16_t c[R];
    uint64_t hb[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long long s = ts + r * 256 + tid;
        c[r] = s < n ? chain[s] : kNoChain;
        hb[r] = __ballot((c[r] & 255) <= (c[r] >> 8));
        if (lane == 0) wpos[r][w] = (uint32_t)__popcll(hb[r]);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int r = 0; r < R; r++)
            for (int ww = 0; ww < W; ww++) {
                const uint32_t v = wpos[r][ww];
                wpos[r][ww] = acc;
                acc += v;
            }
        wpos[0][W] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++)
        if ((hb[r] >> lane) & 1)
            list[wpos[r][w] + (uint32_t)__popcll(hb[r] & below)] = (uint32_t)(r * 256 + tid) << 16 | c[r];
    const int m = (int)wpos[0][W];
    const int Q = (m + 255) / 256;
    __syncthreads();
    for (int q = 0; q < Q; q++) {
        const int e = q * 256 + tid;
        const uint32_t ce = e < m ? list[e] & 0xffffu : kNoChain;
        const int lo = ce & 255, hi = ce >> 8;
        for (int L = 0; L < kMaxLevels; L++) {
            const uint64_t b = __ballot(lo <= L && L <= hi);
            if (lane == 0) pre[q][w][L] = (uint32_t)__popcll(b);
        }
    }
    __syncthreads();
    if (tid < kMaxLevels) {
        uint32_t acc = 0;
        for (int q = 0; q < Q; q++)
            for (int ww = 0; ww < W; ww++) {
                const uint32_t v = pre[q][ww][tid];
                pre[q][ww][tid] = acc;
                acc += v;
            }
    }
    __syncthreads();
    const long long base_end = base[kMaxLevels];  // the node count
    for (int q = 0; q < Q; q++) {
        const int e = q * 256 + tid;
        const uint32_t le = e < m ? list[e] : kNoChain;
        const int u = (int)(le >> 16);
        const long long s = ts + u;
        const int lo = le & 255, hi = (le >> 8) & 255;
        const bool has = lo <= hi;
        const uint64_t key = S.key[u + kTopoPre];
        const int b = has && batch.B > 1 ? upper_index(batch.starOff, batch.B, s) : 0;
        const double half0 = batch.half[b];
        double cx = dec_d(batch.boxEnc[6 * b]) + half0;
        double cy = dec_d(batch.boxEnc[6 * b + 1]) + half0;
        double cz = dec_d(batch.boxEnc[6 * b + 2]) + half0;
        double half = half0;
        // level by level (wave-uniform, so the ballots that rank the star's
        // node in each level are taken by every lane); the node of level L
        // needs the index of level L+1 (its first child)
        auto rank = [&](int L) -> int32_t {
            const uint64_t bl = __ballot(lo <= L && L <= hi);
            return (int32_t)(base[L] + pre[q][w][L] + (uint32_t)__popcll(bl & below));
        };
        int32_t cur = rank(0);
        for (int L = 0; L < kMaxLevels; L++) {
            if (__ballot(has && L <= hi) == 0) break;
            const int32_t nxt = L + 1 < kMaxLevels ? rank(L + 1) : -1;
            if (lo <= L && L <= hi) {
                TravNode t;
                t.x = cx;
                t.y = cy;
                t.z = cz;
                t.crit2 = half;
                t.first = L < hi ? nxt : -1;
                t.next = -1;
                t.lo = (int32_t)s;
                t.hi = -1;
                if (L == hi && L < kMortonBits) {
                    // a leaf of at most leafSize stars: its end is the first
                    // star after s that starts a run of level <= L
                    for (int k = 1; k <= kLeafSize; k++)
                        if (S.delta[u + kTopoPre - 1 + k] < L) {
                            t.hi = (int32_t)(s + k);
                            break;
                        }
                }
                t.nchild = 0;
                t.level = L;
                t.parent = -1;
                t.inst = b;
                if (cur >= 0 && cur < base_end) {  // always (the counts of k_topo_count), checked anyway
                    trav[cur] = t;
                    if (L == lo) chainFirst[s] = cur;
                } else {
                    topo_fail(meta, 1, cur);
                }
            }
            // the child cell holding this star: octant digit L, buildNode's offsets
            const int oct = (int)((key >> (3 * (kMortonBits - 1 - min(L, kMortonBits - 1)))) & 7);
            const double qh = half / 2;
            cx = cx + (((oct >> 2) & 1) ? 1.0 : -1.0) * qh;
            cy = cy + (((oct >> 1) & 1) ? 1.0 : -1.0) * qh;
            cz = cz + ((oct & 1) ? 1.0 : -1.0) * qh;
            half = qh;
            cur = nxt;
        }
    }
}

// End of the level-L run that starts at star lo, for a split node (its run
// holds more than kLeafSize stars).  The stars inside the run start chains at
// levels above L only, and the first star after it starts one at level L or
// below, so the end is the first star from a = lo + kLeafSize + 1 on whose
// chain starts at a level <= L.  It is looked up in the chains of the 64
// stars from a (nine aligned 16-byte loads in flight together, each 32-bit
// word tested for both its stars at once: bit 15 of 0x8000 + L - first is
// set iff first <= L), then, for a longer run, by a galloping search on the
// keys with steps from 64 up.  The chain array has room past n (it lives in
// an 8-byte-per-star buffer); stars at or past end are never taken.
__device__ __forceinline__ int32_t split_end(const uint64_t* __restrict__ keys, const uint16_t* __restrict__ chain,
                                             int32_t lo, int32_t end, int L) {
    const int32_t a = lo + kLeafSize + 1;
    if (a >= end) return end;
    const int32_t a0 = a & ~7;
    uint4 v[9];
    const uint4* p = reinterpret_cast<const uint4*>(chain + a0);
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = p[k];
    const uint32_t c = (0x8000u + (uint32_t)L) * 0x00010001u;
    uint64_t m0 = 0;  // bit i: star a0 + i starts a chain at a level <= L (i < 64)
    uint32_t m1 = 0;  // the same for a0 + 64 + i (i < 8)
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const uint32_t w[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
        for (int h = 0; h < 4; h++) {
            const uint32_t y = (c - (w[h] & 0x00FF00FFu)) & 0x80008000u;
            const uint32_t two = (y >> 15 & 1u) | (y >> 30 & 2u);
            if (k < 8)
                m0 |= (uint64_t)two << (8 * k + 2 * h);
            else
                m1 |= two << (2 * h);
        }
    }
    const int sh = a - a0;  // 0..7: drop the stars before a, keep the window [a, a + 64)
    const uint64_t win = (m0 >> sh) | (sh ? (uint64_t)m1 << (64 - sh) : 0ull);
    if (win) return min(end, a + (int32_t)__builtin_ctzll(win));
    if (a + 64 >= end) return end;
    // the run holds a .. a + 63: gallop on the keys from there
    const int shift = 3 * (kMortonBits - L);
    const uint64_t k0 = keys[lo];
    auto same = [&](int32_t j) { return ((keys[j] ^ k0) >> shift) == 0; };
    int32_t x = a + 63, d = 64;
    while (x + d < end && same(x + d)) {
        x += d;
        d *= 2;
    }
    int32_t y = min(end, x + d);
    while (y - x > 1) {
        const int32_t mid = x + (y - x) / 2;
        if (same(mid))
            x = mid;
        else
            y = mid;
    }
    return y;
}

__global__ __launch_bounds__(256) void k_topo_links(const uint64_t* __restrict__ keys,
                                                    const uint16_t* __restrict__ chain, TravNode* __restrict__ trav,
                                                    const int32_t* __restrict__ chainFirst, Meta* meta,
                                                    Batch batch) {
    if (meta->flags & kFlagNodeOverflow) return;
    const long long nn = meta->nnodes;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nn;
         t += (long long)gridDim.x * blockDim.x) {
        const int32_t lo = trav[t].lo, first = trav[t].first, L = trav[t].level, inst = trav[t].inst;
        if (inst < 0 || inst >= batch.B || L < 0 || L > kMortonBits || lo < batch.starOff[inst] ||
            lo >= batch.starOff[inst + 1]) {  // a record k_topo_nodes did not write (never expected)
            topo_fail(meta, 2, t);
            continue;
        }
        int32_t hi = trav[t].hi;
        const int32_t end = (int32_t)batch.starOff[inst + 1];
        if (hi < 0) hi = L == 0 ? end : split_end(keys, chain, lo, end, L);
        const int32_t next = hi >= end ? -1 : chainFirst[hi];
        // every link must lead to a node that starts where it should (the
        // walks follow them): a first child at lo one level down, a skip link
        // at hi; anything else stops the step before any walk
        int why = hi <= lo || hi > end ? 4 : 0;
        if (next != -1 && (next < 0 || next >= nn || trav[next].lo != hi)) why |= 8;
        if (first >= 0 && L == kMortonBits) why |= 32;
        int nchild = 0;
        if (first >= 0 && !why) {
            const long long cend = meta->levelStart[L + 1] + meta->levelCount[L + 1];
            if (first >= cend || trav[first].lo != lo || trav[first].level != L + 1) why |= 16;
            if (!why) {
                // the children are the level-(L+1) nodes from first that start
                // before hi (at most eight; their starts ascend): all eight
                // starts are loaded at once
                int32_t clo[8];
#pragma unroll
                for (int c = 1; c < 8; c++) clo[c] = first + c < cend ? trav[first + c].lo : INT32_MAX;
                nchild = 1;
#pragma unroll
                for (int c = 1; c < 8; c++) nchild += clo[c] < hi;
                for (int c = 0; c < nchild; c++) trav[first + c].parent = (int32_t)t;
            }
        }
        if (why) topo_fail(meta, why, t);
        trav[t].hi = hi;
        trav[t].next = next;
        trav[t].nchild = nchild;
    }
}

// ---------------------------------------------------------------------------
// moments: finish, build.go:176-217
// ---------------------------------------------------------------------------
__device__ __forceinline__ void quad_add(double q[6], double m, double dx, double dy, double dz) {
    double r2 = dx * dx + dy * dy + dz * dz;
    q[0] += m * (3 * dx * dx - r2);
    q[1] += m * (3 * dy * dy - r2);
    q[2] += m * (3 * dz * dz - r2);
    q[3] += m * 3 * dx * dy;
    q[4] += m * 3 * dx * dz;
    q[5] += m * 3 * dy * dz;
}

// Moments of node ni from its stars (LEAF) or its finished children, in the
// reference's order (finish, build.go:176-217).  While building, the node's
// record holds the cell centre in x,y,z and the half side in crit2.  (Two
// instances, so the leaf kernel carries no registers of the other branch.)
// The stars of a leaf (up to kLeafSize) or the children of a node (up to 8)
// are all loaded into registers before the sums, so that a node waits for one
// round of memory latency instead of one per term; the sums themselves keep
// the reference's order.
template <bool LEAF>
__device__ void node_moments(long long ni, TravNode nd, const double4* __restrict__ spos,
                             TravNode* __restrict__ trav, Mult64* __restrict__ m64, Mult32* __restrict__ m32,
                             Mult16* __restrict__ m16, const double* __restrict__ thetas) {
    const double theta = thetas[nd.inst];
    const double cx = nd.x, cy = nd.y, cz = nd.z, half = nd.crit2;
    double x = 0, y = 0, z = 0, m = 0;
    double q[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (LEAF) {
        // every node holds at least one star; a leaf at the deepest level may
        // hold more than kLeafSize (coincident stars), the rest follow in a loop
        constexpr int P = kLeafSize;
        const int lo = nd.lo, n = nd.hi - nd.lo;
        double4 s[P];
#pragma unroll
        for (int k = 0; k < P; k++) s[k] = spos[lo + min(k, n - 1)];
#pragma unroll
        for (int k = 0; k < P; k++)
            if (k < n) {
                x = x + s[k].w * s[k].x;
                y = y + s[k].w * s[k].y;
                z = z + s[k].w * s[k].z;
                m = m + s[k].w;
            }
        for (int j = lo + P; j < lo + n; j++) {
            double4 p = spos[j];
            x = x + p.w * p.x;
            y = y + p.w * p.y;
            z = z + p.w * p.z;
            m = m + p.w;
        }
        if (m > 0) {
            x = x / m;
            y = y / m;
            z = z / m;
        } else {
            x = cx;
            y = cy;
            z = cz;
        }
#pragma unroll
        for (int k = 0; k < P; k++)
            if (k < n) quad_add(q, s[k].w, s[k].x - x, s[k].y - y, s[k].z - z);
        for (int j = lo + P; j < lo + n; j++) {
            double4 p = spos[j];
            quad_add(q, p.w, p.x - x, p.y - y, p.z - z);
        }
    } else {
        const int first = nd.first, n = nd.nchild;  // 1 .. 8
        Mult64 cm[8];
        double4 cp[8];  // the children's centres of mass (x, y, z; w unused)
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int c = first + min(k, n - 1);
            cm[k] = m64[c];
            cp[k] = *(const double4*)&trav[c];
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (k < n) {
                x = x + cm[k].m * cp[k].x;
                y = y + cm[k].m * cp[k].y;
                z = z + cm[k].m * cp[k].z;
                m = m + cm[k].m;
            }
        if (m > 0) {
            x = x / m;
            y = y / m;
            z = z / m;
        } else {
            x = cx;
            y = cy;
            z = cz;
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (k < n) {
                q[0] = q[0] + cm[k].qxx;
                q[1] = q[1] + cm[k].qyy;
                q[2] = q[2] + cm[k].qzz;
                q[3] = q[3] + cm[k].qxy;
                q[4] = q[4] + cm[k].qxz;
                q[5] = q[5] + cm[k].qyz;
                quad_add(q, cm[k].m, cp[k].x - x, cp[k].y - y, cp[k].z - z);
            }
    }
    double delta = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    double crit = 2 * half / theta + delta;
    Mult64 mm;
    mm.m = m;
    mm.qxx = q[0]; mm.qyy = q[1]; mm.qzz = q[2];
    mm.qxy = q[3]; mm.qxz = q[4]; mm.qyz = q[5];
    mm.pad = half;  // the cell's half side (scale of the FP16-stored quadrupoles)
    m64[ni] = mm;
    Mult32 mf;
    mf.m = (float)m;
    mf.qxx = (float)q[0]; mf.qyy = (float)q[1]; mf.qzz = (float)q[2];
    mf.qxy = (float)q[3]; mf.qxz = (float)q[4]; mf.qyz = (float)q[5];
    mf.pad = 0;
    m32[ni] = mf;
    Mult16 mh;
    mh.m = (float)m;
    const double inv = m > 0 ? 1.0 / (m * half * half) : 0.0;
    for (int c = 0; c < 6; c++) mh.q[c] = __half_as_ushort(__float2half_rn((float)(q[c] * inv)));
    m16[ni] = mh;
    nd.x = x;
    nd.y = y;
    nd.z = z;
    nd.crit2 = crit * crit;
    trav[ni] = nd;
}

// Internal nodes of level L (their children, one level down, are done).
// (Finishing two levels per launch, each thread doing its node's split
// children first, measured 2.6x slower: the per-thread chains of dependent
// loads cost more than the launches they save.)
__global__ __launch_bounds__(256) void k_moments_level(int L, const double4* __restrict__ spos,
                                                       TravNode* __restrict__ trav, Mult64* __restrict__ m64,
                                                       Mult32* __restrict__ m32, Mult16* __restrict__ m16,
                                                       const Meta* __restrict__ meta, Batch batch) {
    const long long start = meta->levelStart[L];
    const long long cnt = meta->levelCount[L];
    if (meta->flags & kFlagTreeIncomplete) return;
    // node t of the level goes to thread t / gridDim.x of workgroup
    // t % gridDim.x: a level's internal nodes (a few hundred to a few
    // thousand) are spread over every CU instead of packed into the first
    // workgroups, so more of their gathers are in flight at once
    for (long long t = (long long)threadIdx.x * gridDim.x + blockIdx.x; t < cnt;
         t += (long long)gridDim.x * blockDim.x) {
        const TravNode nd = trav[start + t];
        if (nd.first >= 0) node_moments<false>(start + t, nd, spos, trav, m64, m32, m16, batch.theta);
    }
}

// Every leaf of the tree, whatever its level, in one launch: a leaf's moments
// depend only on its stars, so the level-by-level launches that follow only
// carry the internal nodes (short sums over at most eight children) and no
// longer wait, level after level, on the leaves' sums over up to 24 stars.
// The nodes of all levels are numbered contiguously from 0.
// It also counts the nodes (meta->nnodes) and checks that no split node lies
// at level Lmax or below, which the per-level launches (levels Lmax-1 .. 0)
// would not reach (a deeper tree raises kFlagDepth and the step is repeated
// with every level).
__global__ __launch_bounds__(256) void k_moments_leaves(const double4* __restrict__ spos, TravNode* __restrict__ trav,
                                                        Mult64* __restrict__ m64, Mult32* __restrict__ m32,
                                                        Mult16* __restrict__ m16, Meta* __restrict__ meta,
                                                        Batch batch, int Lmax) {
    if (meta->flags & kFlagTreeIncomplete) return;
    long long cnt = 0;
    for (int l = 0; l <= kMaxLevels; l++) cnt += meta->levelCount[l];
    if (blockIdx.x == 0 && threadIdx.x == 0) meta->nnodes = cnt;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < cnt;
         t += (long long)gridDim.x * blockDim.x) {
        const TravNode nd = trav[t];
        if (nd.first < 0)
            node_moments<true>(t, nd, spos, trav, m64, m32, m16, batch.theta);
        else if (nd.level >= Lmax)
            atomicOr(&meta->flags, (int)kFlagDepth);
    }
}


// Zeroes the holes' accelerations (bh.go:87-89, 99-101); with accumulate the
// step's flags also go into the sticky word (the last kernel of a step).
__global__ void k_zero_holes(const int32_t* __restrict__ holes, Meta* __restrict__ meta, double* ax,
                             double* ay, double* az, int accumulate) {
    long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (accumulate && h == 0 && meta->flags) atomicOr(&meta->sticky, (unsigned long long)meta->flags);
    for (; h < meta->nholes; h += (long long)gridDim.x * blockDim.x) {
        int32_t i = holes[h];
        ax[i] = 0;
        ay[i] = 0;
        az[i] = 0;
    }
}

}  // namespace gx

namespace gx {

// ---------------------------------------------------------------------------
// Active sets.  A force evaluation with an active mask (state order) builds
// the tree from all stars as usual but walks only the active ones: their
// sorted indices are compacted, in sorted order, into the target list, and
// each instance's walk groups are cut from its own run of targets.  Forces of
// inactive bodies are left as they are.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_active_count(const uint8_t* __restrict__ active,
                                                      const int32_t* __restrict__ order, const Meta* __restrict__ meta,
                                                      uint8_t* __restrict__ actS, int* __restrict__ blockAct) {
    __shared__ int wsum[4];
    const long long nst = meta->nstars;
    const int tid = threadIdx.x;
    constexpr int per = kCollectTile / 256;
    const long long base = (long long)blockIdx.x * kCollectTile;
    int cnt = 0;
    for (int j = 0; j < per; j++) {
        const long long k = base + (long long)j * 256 + tid;  // coalesced over the tile
        if (k < nst) {
            const uint8_t a = active[order[k]] != 0;
            actS[k] = a;
            cnt += a;
        }
    }
    cnt = wave_sum_i(cnt);
    if (lane_id() == 0) wsum[tid / kWave] = cnt;
    __syncthreads();
    if (tid == 0) blockAct[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// One workgroup: exclusive scan of the tile counts; the target count.
__global__ __launch_bounds__(1024) void k_active_scan(int nb, int* __restrict__ blockAct, Meta* meta) {
    __shared__ long long ws[1][kScanThreads / kWave + 1];
    const int tid = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int lo = tid * per, hi = min(nb, lo + per);
    long long v[1] = {0}, total[1];
    for (int i = lo; i < hi; i++) v[0] += blockAct[i];
    block_scan_1024<1>(v, total, ws);
    if (tid == 0) meta->ntargets = total[0];
    long long run = v[0];
    for (int i = lo; i < hi; i++) {
        const int c = blockAct[i];
        blockAct[i] = (int)run;
        run += c;
    }
}

// One wavefront per instance boundary b in [0, B]: tOff[b] = active stars
// before sorted position starOff[b] (tile prefix + count inside the tile).
__global__ __launch_bounds__(256) void k_active_inst(const uint8_t* __restrict__ actS,
                                                     const int* __restrict__ blockAct, const Meta* __restrict__ meta,
                                                     Batch batch, long long* __restrict__ tOff) {
    const int b = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (b > batch.B) return;
    const long long pos = batch.starOff[b];
    if (pos >= meta->nstars) {
        if (lane_id() == 0) tOff[b] = meta->ntargets;
        return;
    }
    const long long tile = pos / kCollectTile;
    int cnt = 0;
    for (long long k = tile * kCollectTile + lane_id(); k < pos; k += kWave) cnt += actS[k];
    cnt = wave_sum_i(cnt);
    if (lane_id() == 0) tOff[b] = blockAct[tile] + cnt;
}

// One workgroup: walk groups per instance from its target count.
__global__ __launch_bounds__(1024) void k_active_groups(Batch batch, const long long* __restrict__ tOff) {
    __shared__ long long ws[1][kScanThreads / kWave + 1];
    const int tid = threadIdx.x;
    const int B = batch.B;
    const int per = (B + 1023) / 1024;
    const int lo = tid * per, hi = min(B, lo + per);
    long long v[1] = {0}, total[1];
    for (int b = lo; b < hi; b++) v[0] += (tOff[b + 1] - tOff[b] + batch.gsize - 1) / batch.gsize;
    block_scan_1024<1>(v, total, ws);
    if (tid == 0) batch.groupOff[B] = total[0];
    long long rg = v[0];
    for (int b = lo; b < hi; b++) {
        batch.groupOff[b] = rg;
        rg += (tOff[b + 1] - tOff[b] + batch.gsize - 1) / batch.gsize;
    }
}

// Stable compaction of the active sorted positions into the target list.
__global__ __launch_bounds__(256) void k_active_scatter(const uint8_t* __restrict__ actS,
                                                        const int* __restrict__ blockAct,
                                                        const Meta* __restrict__ meta, int32_t* __restrict__ tlist) {
    __shared__ int ssum[256];
    const long long nst = meta->nstars;
    const int tid = threadIdx.x;
    constexpr int per = kCollectTile / 256;
    const long long base = (long long)blockIdx.x * kCollectTile + (long long)tid * per;
    uint32_t amask = 0;
    for (int j = 0; j < per; j++) {
        const long long k = base + j;
        if (k < nst && actS[k]) amask |= 1u << j;
    }
    const int c = __popc(amask);
    ssum[tid] = c;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int a = tid >= d ? ssum[tid - d] : 0;
        __syncthreads();
        ssum[tid] += a;
        __syncthreads();
    }
    int o = blockAct[blockIdx.x] + ssum[tid] - c;
    for (int j = 0; j < per; j++)
        if (amask >> j & 1) tlist[o++] = (int32_t)(base + j);
}

}  // namespace gx

namespace gx {

// ---------------------------------------------------------------------------
// KDK leapfrog pieces: kick and drift, internal/sim/sim.go:302-343 (the
// Newtonian branch, c <= 0), over every body as in the reference.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kick(long long n, double* __restrict__ vx, double* __restrict__ vy,
                                              double* __restrict__ vz, const double* __restrict__ ax,
                                              const double* __restrict__ ay, const double* __restrict__ az, double h) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        vx[i] += h * ax[i];
        vy[i] += h * ay[i];
        vz[i] += h * az[i];
    }
}

__global__ __launch_bounds__(256) void k_drift(long long n, double* __restrict__ px, double* __restrict__ py,
                                               double* __restrict__ pz, const double* __restrict__ vx,
                                               const double* __restrict__ vy, const double* __restrict__ vz,
                                               double dt) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        px[i] += dt * vx[i];
        py[i] += dt * vy[i];
        pz[i] += dt * vz[i];
    }
}

// ---------------------------------------------------------------------------
// Block time steps with power-of-two rungs for gravity.  The rung arithmetic
// is the reference's for gas (internal/sim/blockstep.go:23-45, 81-90): rung r
// means a step of T/2^r (maxRung = levels - 1 caps it), time is counted in
// ticks of T/2^maxRung, and a step may start only at a multiple of its own
// length.  A body is therefore active at tick t exactly when t is a multiple
// of its step length.  The wanted step of a star is sqrt(2 eta eps / |a|)
// (the acceleration criterion of GADGET-2, Springel 2005, eq. 34); the
// reference has no gravity step criterion of its own.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rung_for(double dt, double T, int maxRung) {  // blockstep.go:37-45
    int r = 0;
    double x = T;
    while (x > dt && r < maxRung) {
        x *= 0.5;
        r++;
    }
    return r;
}

// active[i] = body i's step ends at tick now.
__global__ __launch_bounds__(256) void k_block_active(long long n, const int8_t* __restrict__ rung, int now, int full,
                                                      uint8_t* __restrict__ active) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        active[i] = now % (full >> rung[i]) == 0;
}

// At tick now: the closing kick of every body whose step ends there (now > 0),
// then, before the end of the global step, its new rung (aligned to now, as
// startStep, blockstep.go:81-90) and the opening kick of the new step.  With
// now = 0 every body starts a step.  hist[r] counts the bodies on rung r
// afterwards (all bodies, active or not).
__global__ __launch_bounds__(256) void k_block_kick(long long n, double* __restrict__ vx, double* __restrict__ vy,
                                                    double* __restrict__ vz, const double* __restrict__ ax,
                                                    const double* __restrict__ ay, const double* __restrict__ az,
                                                    const uint8_t* __restrict__ kind, const uint8_t* __restrict__ alive,
                                                    int8_t* __restrict__ rung, int now, int full, int maxRung, double T,
                                                    double crit, unsigned long long* __restrict__ hist) {
    __shared__ unsigned int bh[4][8];
    unsigned int cnt[8] = {};  // rung populations of this thread's bodies (maxRung <= 7)
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int r = rung[i];
        if (now == 0 || now % (full >> r) == 0) {
            if (now > 0) {
                const double h = 0.5 * (T / (double)(1 << r));
                vx[i] += h * ax[i];
                vy[i] += h * ay[i];
                vz[i] += h * az[i];
            }
            if (now < full) {
                r = 0;
                if (alive[i] && kind[i] != kHole) {
                    const double x = ax[i], y = ay[i], z = az[i];
                    const double a = sqrt(x * x + y * y + z * z);
                    r = rung_for(sqrt(crit / a), T, maxRung);
                }
                while (now % (full >> r) != 0) r++;
                rung[i] = (int8_t)r;
                const double h = 0.5 * (T / (double)(1 << r));
                vx[i] += h * ax[i];
                vy[i] += h * ay[i];
                vz[i] += h * az[i];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q++) cnt[q] += r == q;
    }
    // block totals, then one atomic per populated rung and workgroup
    const int w = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int c = wave_sum_i((int)cnt[q]);
        if (lane_id() == 0) bh[w][q] = (unsigned int)c;
    }
    __syncthreads();
    if (threadIdx.x <= maxRung) {
        const int q = threadIdx.x;
        const unsigned long long c = (unsigned long long)bh[0][q] + bh[1][q] + bh[2][q] + bh[3][q];
        if (c) atomicAdd(&hist[q], c);
    }
}

// p[0 .. count) = v, for any count and launch shape (grid-stride: a single
// 64-thread block also fills a Meta of more than 64 words).
__global__ void k_fill_u64(unsigned long long* p, int count, unsigned long long v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = v;
}

// The walk's counters and flags, before a walk that repeats one on the same
// tree (the autotuner's timed walks).
__global__ void k_walk_reset(Meta* meta) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    meta->pp = meta->pc = 0;
    for (int q = 0; q < 8; q++) meta->xcdNext[q] = 0;
    meta->listEntries = meta->rounds = meta->mixed = 0;
    meta->maxStack = 0;
    meta->maxList[0] = meta->maxList[1] = meta->maxList[2] = 0;
    meta->flags &= ~(kFlagStackOverflow | kFlagListOverflow);
}

// ---------------------------------------------------------------------------
// FP64 diagnostics, internal/sim/diag.go: per workgroup partial sums, in a
// fixed order, of
//   0      kinetic energy of the sources, m u^2 / (1 + 1)  (sourceKinetic :68-77,
//          kinetic relativity.go:24-26 with c <= 0)
//   1..3   momentum m v                                  (Momentum :102-109)
//   4..6   angular momentum x cross (m v)                (AngularMomentum :112-119)
//   7      mass, 8..10 mass-weighted position            (CenterOfMass :122-134)
// over every body, as the reference (Alive is not consulted there either).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag(long long n, const double* __restrict__ px,
                                              const double* __restrict__ py, const double* __restrict__ pz,
                                              const double* __restrict__ vx, const double* __restrict__ vy,
                                              const double* __restrict__ vz, const double* __restrict__ m,
                                              const uint8_t* __restrict__ kind, int dmOff,
                                              double* __restrict__ partial) {
    __shared__ double red[kDiagValues][256];
    double acc[kDiagValues] = {};
    const int tid = threadIdx.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + tid; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double mi = m[i], ux = vx[i], uy = vy[i], uz = vz[i];
        const double x = px[i], y = py[i], z = pz[i];
        if (!(dmOff && kind[i] == kDark)) acc[0] += mi * (ux * ux + uy * uy + uz * uz) / (1.0 + 1.0);
        const double qx = ux * mi, qy = uy * mi, qz = uz * mi;
        acc[1] += qx;
        acc[2] += qy;
        acc[3] += qz;
        acc[4] += y * qz - z * qy;
        acc[5] += z * qx - x * qz;
        acc[6] += x * qy - y * qx;
        acc[7] += mi;
        acc[8] += x * mi;
        acc[9] += y * mi;
        acc[10] += z * mi;
    }
    for (int c = 0; c < kDiagValues; c++) red[c][tid] = acc[c];
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (tid < d)
            for (int c = 0; c < kDiagValues; c++) red[c][tid] += red[c][tid + d];
        __syncthreads();
    }
    if (tid < kDiagValues) partial[(size_t)blockIdx.x * kDiagValues + tid] = red[tid][0];
}

}  // namespace gx
