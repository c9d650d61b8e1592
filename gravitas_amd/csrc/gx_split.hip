// gx_split.hip -- multiple time stepping near black holes on the GPU,
// internal/sim/split.go (an impulse/RESPA scheme): star-star forces come from
// the tree once per step; every force involving a hole is a direct sum here.
// Stars whose timescale at a hole is below 20 outer steps ("near") and all
// holes take power-of-two substeps; far stars receive hole forces at the
// outer kicks.  Restated for states without gas; with c > 0 the pairs with
// holes use the Paczynski-Wiita radius and drifts use u / gamma
// (relativity.go), and in the retarded gravity modes (retarded.go) a body
// feels each hole where it was on the body's past light cone, read from the
// hole's history of outer steps.
//
// Compiled with -ffp-contract=off: per-body sums over the holes run in the
// hole order of the reference and are bit-identical to it; sums over bodies
// for a hole are tree reductions in a fixed order (the reference adds chunk
// results in completion order, so it fixes no order either).
#include "gx_common.h"
#include "gx_kernels.h"
#include "gx_pair.h"

namespace gx {

// holeHistory.at, retarded.go:85-103: the hole's position at time t,
// interpolated linearly between the two outer steps around t, or extrapolated
// from the oldest entry before it; the velocity too if vel is non-null.
// False if the hole had not appeared yet at t.
__device__ __forceinline__ bool hist_at(const HistDev& h, double t, double* pos, double* vel) {
    if (t < h.appeared) return false;
    const double* rx = h.ring;
    const double* ry = rx + kHistLen;
    const double* rz = ry + kHistLen;
    const double* rvx = rz + kHistLen;
    const double* rvy = rvx + kHistLen;
    const double* rvz = rvy + kHistLen;
    const int oldest = (h.head - h.n + 1 + kHistLen) % kHistLen;
    if (t <= h.t0 || h.n == 1) {
        const double d = t - h.t0;
        const double ux = rvx[oldest], uy = rvy[oldest], uz = rvz[oldest];
        pos[0] = rx[oldest] + d * ux;
        pos[1] = ry[oldest] + d * uy;
        pos[2] = rz[oldest] + d * uz;
        if (vel) {
            vel[0] = ux;
            vel[1] = uy;
            vel[2] = uz;
        }
        return true;
    }
    const double f = (t - h.t0) / h.dt;
    int k = (int)f;
    if (k > h.n - 2) k = h.n - 2;
    const double w = f - (double)k;
    const int a = (oldest + k) % kHistLen, b = (oldest + k + 1) % kHistLen;
    pos[0] = rx[a] + w * (rx[b] - rx[a]);
    pos[1] = ry[a] + w * (ry[b] - ry[a]);
    pos[2] = rz[a] + w * (rz[b] - rz[a]);
    if (vel) {
        vel[0] = rvx[a] + w * (rvx[b] - rvx[a]);
        vel[1] = rvy[a] + w * (rvy[b] - rvy[a]);
        vel[2] = rvz[a] + w * (rvz[b] - rvz[a]);
    }
    return true;
}

// sourcePos, retarded.go:118-155, in a retarded mode (r.hist non-null): where
// a body at (x, y, z) sees hole slot k, whose current position (hx, hy, hz)
// is replaced by the source; false if its gravity has not arrived yet.  The
// light-cone delay is a fixed point iterated to 1e-7 time units (at most 64
// times); the extrapolated mode adds delay * v.
__device__ __forceinline__ bool source_pos(const RetardDev& r, long long k, double x, double y, double z, double c,
                                        double& hx, double& hy, double& hz) {
    const HistDev h = r.hist[k];
    double delay = sqrt((x - hx) * (x - hx) + (y - hy) * (y - hy) + (z - hz) * (z - hz)) / c;
    if (!h.ring || delay < r.dt) return true;  // within one step: below the step resolution
    double p[3], v[3];
    for (int it = 0; it < kRetardMaxIt; it++) {
        if (!hist_at(h, r.time - delay, p, nullptr)) {
            hx = hy = hz = 0;
            return false;
        }
        const double next = sqrt((x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) + (z - p[2]) * (z - p[2])) / c;
        const bool converged = fabs(next - delay) < kRetardTol;
        delay = next;
        if (converged) break;
    }
    hist_at(h, r.time - delay, p, v);
    if (r.mode == 2) {
        p[0] = p[0] + delay * v[0];
        p[1] = p[1] + delay * v[1];
        p[2] = p[2] + delay * v[2];
    }
    hx = p[0];
    hy = p[1];
    hz = p[2];
    return true;
}

// holePull, split.go:120-137: pull of every hole except i itself on body i,
// from where the body sees it (sourcePos; the hole itself with instant
// gravity).
__device__ __forceinline__ void hole_pull(long long i, const double* __restrict__ x, const double* __restrict__ y,
                                          const double* __restrict__ z, const double* __restrict__ mass,
                                          const uint8_t* __restrict__ kind, const int32_t* __restrict__ holes,
                                          long long nh, double eps2, double c, const RetardDev& r, double& ax,
                                          double& ay, double& az) {
    ax = ay = az = 0;
    const double xi = x[i], yi = y[i], zi = z[i];
    const int ki = kind[i];
    const double mi = mass[i];
    for (long long k = 0; k < nh; k++) {
        const int32_t h = holes[k];
        if (h == i) continue;
        double hx = x[h], hy = y[h], hz = z[h];
        if (r.hist && !source_pos(r, k, xi, yi, zi, c, hx, hy, hz)) continue;  // not arrived yet
        const double dx = hx - xi, dy = hy - yi, dz = hz - zi;
        const double rs = pair_rs(ki, mi, kHole, mass[h], c);
        const double f = pair_force(mass[h], dx * dx + dy * dy + dz * dz, eps2, rs);
        ax = ax + f * dx;
        ay = ay + f * dy;
        az = az + f * dz;
    }
}

// holeTimescale, split.go:27-36.
__device__ __forceinline__ double hole_timescale(long long i, int32_t h, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ mass, const uint8_t* __restrict__ kind,
                                                 double eps2, double c) {
    const double dx = x[i] - x[h], dy = y[i] - y[h], dz = z[i] - z[h];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rs = pair_rs(kind[i], mass[i], kHole, mass[h], c);
    const double a = pair_force(mass[h], r2, eps2, rs) * sqrt(r2);
    if (a == 0) return INFINITY;
    return sqrt(sqrt(r2 + eps2) / a);
}

// classify, split.go:39-80: near flags and the shortest timescale below the
// limit (atomic minimum on the order-preserving encoding of doubles).
__global__ __launch_bounds__(256) void k_split_classify(long long n, const double* __restrict__ x,
                                                        const double* __restrict__ y, const double* __restrict__ z,
                                                        const double* __restrict__ mass,
                                                        const uint8_t* __restrict__ kind,
                                                        const int32_t* __restrict__ holes,
                                                        const Meta* __restrict__ meta, double eps2, double c,
                                                        double limit, uint8_t* __restrict__ nearFlag,
                                                        unsigned long long* __restrict__ tauEnc) {
    const long long nh = meta->nholes;
    double local = INFINITY;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        uint8_t near = 0;
        if (kind[i] != kHole) {
            for (long long k = 0; k < nh; k++) {
                const double tau = hole_timescale(i, holes[k], x, y, z, mass, kind, eps2, c);
                if (tau < limit) {
                    near = 1;
                    local = fmin(local, tau);
                }
            }
        }
        nearFlag[i] = near;
    }
    local = wave_min_d(local);
    if (lane_id() == 0 && local < INFINITY) atomicMin(tauEnc, enc_d(local));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double hm = INFINITY;
        for (long long a = 0; a < nh; a++)
            for (long long b = a + 1; b < nh; b++) {
                const double tau = hole_timescale(holes[a], holes[b], x, y, z, mass, kind, eps2, c);
                if (tau < limit) hm = fmin(hm, tau);
            }
        if (hm < INFINITY) atomicMin(tauEnc, enc_d(hm));
    }
}

// Stable compaction of nonzero flags (state order): per-tile counts, one
// scan, scatter.  Tiles of kCollectTile bodies.
__global__ __launch_bounds__(256) void k_flag_count(long long n, const uint8_t* __restrict__ flag,
                                                    int* __restrict__ tileCnt) {
    __shared__ int wsum[4];
    int cnt = 0;
    const long long base = (long long)blockIdx.x * kCollectTile;
    for (int j = 0; j < kCollectTile / 256; j++) {
        const long long i = base + (long long)j * 256 + threadIdx.x;
        if (i < n) cnt += flag[i] != 0;
    }
    cnt = wave_sum_i(cnt);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) tileCnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(1024) void k_flag_scan(int nb, int* __restrict__ tileCnt, long long* __restrict__ total) {
    __shared__ long long ss[1024];
    const int tid = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int lo = tid * per, hi = min(nb, lo + per);
    long long s0 = 0;
    for (int i = lo; i < hi; i++) s0 += tileCnt[i];
    ss[tid] = s0;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const long long a = tid >= d ? ss[tid - d] : 0;
        __syncthreads();
        ss[tid] += a;
        __syncthreads();
    }
    long long run = ss[tid] - s0;
    if (tid == 1023) *total = ss[1023];
    for (int i = lo; i < hi; i++) {
        const int cc = tileCnt[i];
        tileCnt[i] = (int)run;
        run += cc;
    }
}

__global__ __launch_bounds__(256) void k_flag_scatter(long long n, const uint8_t* __restrict__ flag,
                                                      const int* __restrict__ tileOff, int32_t* __restrict__ list) {
    __shared__ int ssum[256];
    const int tid = threadIdx.x;
    constexpr int per = kCollectTile / 256;
    const long long base = (long long)blockIdx.x * kCollectTile + (long long)tid * per;
    uint32_t m = 0;
    for (int j = 0; j < per; j++)
        if (base + j < n && flag[base + j]) m |= 1u << j;
    const int cnt = __popc(m);
    ssum[tid] = cnt;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int a = tid >= d ? ssum[tid - d] : 0;
        __syncthreads();
        ssum[tid] += a;
        __syncthreads();
    }
    int o = tileOff[blockIdx.x] + ssum[tid] - cnt;
    for (int j = 0; j < per; j++)
        if (m >> j & 1) list[o++] = (int32_t)(base + j);
}

// outerHoleForces, split.go:82-117, the bodies' side: far stars get the pull
// of every hole, near stars and holes 0 (the holes' side is added by
// k_split_hole_sum).
__global__ __launch_bounds__(256) void k_split_far(long long n, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   const double* __restrict__ mass, const uint8_t* __restrict__ kind,
                                                   const uint8_t* __restrict__ nearFlag,
                                                   const int32_t* __restrict__ holes, const Meta* __restrict__ meta,
                                                   double eps2, double c, RetardDev r, double* __restrict__ fx,
                                                   double* __restrict__ fy, double* __restrict__ fz) {
    const long long nh = meta->nholes;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double ax = 0, ay = 0, az = 0;
        if (kind[i] != kHole && !nearFlag[i]) hole_pull(i, x, y, z, mass, kind, holes, nh, eps2, c, r, ax, ay, az);
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
    }
}

// fastForces, split.go:141-160, the near stars' side: the pull of every hole.
__global__ __launch_bounds__(256) void k_split_fast(const int32_t* __restrict__ nearList, const long long* __restrict__ nn,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ mass,
                                                    const uint8_t* __restrict__ kind, const int32_t* __restrict__ holes,
                                                    const Meta* __restrict__ meta, double eps2, double c,
                                                    RetardDev r, double* __restrict__ fx, double* __restrict__ fy,
                                                    double* __restrict__ fz) {
    const long long nh = meta->nholes, cnt = *nn;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += (long long)gridDim.x * blockDim.x) {
        const int32_t i = nearList[j];
        double ax, ay, az;
        hole_pull(i, x, y, z, mass, kind, holes, nh, eps2, c, r, ax, ay, az);
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
    }
}

// The holes' side of outerHoleForces (far: every non-hole body that is not
// near) and of fastForces (near: the near list, after the pull of the other
// holes).  Workgroup (k, j) sums slice j of the bodies for hole k into
// partial[(k * slices + j) * 3 ..]; k_split_hole_finish adds the slices in
// order.  Both sums are fixed-order tree reductions.
__global__ __launch_bounds__(256) void k_split_hole_sum(long long n, const double* __restrict__ x,
                                                        const double* __restrict__ y, const double* __restrict__ z,
                                                        const double* __restrict__ mass,
                                                        const uint8_t* __restrict__ kind, int dmOff,
                                                        const uint8_t* __restrict__ nearFlag,
                                                        const int32_t* __restrict__ nearList,
                                                        const long long* __restrict__ nn,
                                                        const int32_t* __restrict__ holes,
                                                        const Meta* __restrict__ meta, double eps2, double c,
                                                        int nearMode, double* __restrict__ partial) {
    __shared__ double red[3][256];
    const long long k = blockIdx.x;
    const long long nh = meta->nholes;
    if (k >= nh) return;
    const int32_t h = holes[k];
    const double hx = x[h], hy = y[h], hz = z[h], rs = schwarzschild(mass[h], c);
    const long long cnt = nearMode ? *nn : n;
    const long long per = (cnt + gridDim.y - 1) / gridDim.y;
    const long long lo = (long long)blockIdx.y * per, hi = min(cnt, lo + per);
    double ax = 0, ay = 0, az = 0;
    for (long long j = lo + threadIdx.x; j < hi; j += blockDim.x) {
        long long i = j;
        if (nearMode)
            i = nearList[j];
        else if (kind[i] == kHole || nearFlag[i])
            continue;
        const double dx = x[i] - hx, dy = y[i] - hy, dz = z[i] - hz;
        const double sm = (dmOff && kind[i] == kDark) ? 0.0 : mass[i];  // SourceMass, sim.go:93-98
        const double f = pair_force(sm, dx * dx + dy * dy + dz * dz, eps2, rs);
        ax = ax + f * dx;
        ay = ay + f * dy;
        az = az + f * dz;
    }
    red[0][threadIdx.x] = ax;
    red[1][threadIdx.x] = ay;
    red[2][threadIdx.x] = az;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[(k * gridDim.y + blockIdx.y) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_split_hole_finish(const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, const double* __restrict__ mass,
                                    const uint8_t* __restrict__ kind, const int32_t* __restrict__ holes,
                                    const Meta* __restrict__ meta, double eps2, double c, RetardDev r, int nearMode,
                                    int slices, const double* __restrict__ partial, double* __restrict__ fx,
                                    double* __restrict__ fy, double* __restrict__ fz) {
    const long long nh = meta->nholes;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nh; k += (long long)gridDim.x * blockDim.x) {
        const int32_t h = holes[k];
        double ax = 0, ay = 0, az = 0;
        if (nearMode) hole_pull(h, x, y, z, mass, kind, holes, nh, eps2, c, r, ax, ay, az);
        double sx = 0, sy = 0, sz = 0;
        for (int j = 0; j < slices; j++) {
            sx += partial[(k * slices + j) * 3];
            sy += partial[(k * slices + j) * 3 + 1];
            sz += partial[(k * slices + j) * 3 + 2];
        }
        fx[h] = ax + sx;
        fy[h] = ay + sy;
        fz[h] = az + sz;
    }
}

// kickOuter, split.go:201-208: v += h (a + far) over every body.
__global__ __launch_bounds__(256) void k_split_kick_outer(long long n, double* __restrict__ vx, double* __restrict__ vy,
                                                          double* __restrict__ vz, const double* __restrict__ ax,
                                                          const double* __restrict__ ay, const double* __restrict__ az,
                                                          const double* __restrict__ fx, const double* __restrict__ fy,
                                                          const double* __restrict__ fz, double h) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        vx[i] += h * (ax[i] + fx[i]);
        vy[i] += h * (ay[i] + fy[i]);
        vz[i] += h * (az[i] + fz[i]);
    }
}

// kickFast (split.go:210-219) and driftIdx (:221-229) over a list of bodies;
// the list length is either given or read from *cntDev.
__global__ __launch_bounds__(256) void k_split_kick_list(const int32_t* __restrict__ list, long long cntHost,
                                                         const long long* __restrict__ cntDev, double* __restrict__ vx,
                                                         double* __restrict__ vy, double* __restrict__ vz,
                                                         const double* __restrict__ fx, const double* __restrict__ fy,
                                                         const double* __restrict__ fz, double h) {
    const long long cnt = cntDev ? *cntDev : cntHost;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += (long long)gridDim.x * blockDim.x) {
        const int32_t i = list[j];
        vx[i] += h * fx[i];
        vy[i] += h * fy[i];
        vz[i] += h * fz[i];
    }
}

__global__ __launch_bounds__(256) void k_split_drift_list(const int32_t* __restrict__ list, long long cntHost,
                                                          const long long* __restrict__ cntDev, double* __restrict__ x,
                                                          double* __restrict__ y, double* __restrict__ z,
                                                          const double* __restrict__ vx, const double* __restrict__ vy,
                                                          const double* __restrict__ vz, double dt, double c) {
    const long long cnt = cntDev ? *cntDev : cntHost;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += (long long)gridDim.x * blockDim.x) {
        const int32_t i = list[j];
        const double f = dt / gamma_of(vx[i], vy[i], vz[i], c);
        x[i] += f * vx[i];
        y[i] += f * vy[i];
        z[i] += f * vz[i];
    }
}

// driftFar, split.go:225-238 (no gas): every body that is neither a hole nor
// near.  With nearFlag null it is drift (sim.go:316-343) over every body.
__global__ __launch_bounds__(256) void k_split_drift_far(long long n, const uint8_t* __restrict__ kind,
                                                         const uint8_t* __restrict__ nearFlag, double* __restrict__ x,
                                                         double* __restrict__ y, double* __restrict__ z,
                                                         const double* __restrict__ vx, const double* __restrict__ vy,
                                                         const double* __restrict__ vz, double dt, double c) {
    const double invC2 = c > 0 ? 1 / (c * c) : 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        if (nearFlag && (kind[i] == kHole || nearFlag[i])) continue;
        const double ux = vx[i], uy = vy[i], uz = vz[i];
        double f;
        if (nearFlag)
            f = dt / gamma_of(ux, uy, uz, c);  // driftFar
        else if (c > 0)
            f = dt / sqrt(1 + (ux * ux + uy * uy + uz * uz) * invC2);  // drift, sim.go:334-341
        else
            f = dt;  // drift, sim.go:325-331
        x[i] += f * ux;
        y[i] += f * uy;
        z[i] += f * uz;
    }
}

// holeHistory.push, retarded.go:75-83 (the ring part; the host keeps head, n
// and t0): the position and coordinate velocity u / gamma
// (CoordinateVelocity, relativity.go:29-32) of each listed body into its slot.
__global__ __launch_bounds__(64) void k_retard_push(const HistPush* __restrict__ p, int cnt,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ vx,
                                                    const double* __restrict__ vy, const double* __restrict__ vz,
                                                    double c) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
        const HistPush q = p[j];
        const int32_t b = q.body;
        const double ux = vx[b], uy = vy[b], uz = vz[b];
        const double ig = 1 / gamma_of(ux, uy, uz, c);
        double* ring = q.ring + q.pos;
        ring[0] = x[b];
        ring[kHistLen] = y[b];
        ring[2 * kHistLen] = z[b];
        ring[3 * kHistLen] = ux * ig;
        ring[4 * kHistLen] = uy * ig;
        ring[5 * kHistLen] = uz * ig;
    }
}

// sourcePos of hole slot k for m observers (the test hook of the retarded
// modes): (sx, sy, sz) and visible per observer.
__global__ __launch_bounds__(256) void k_retard_source(RetardDev r, long long k, const int32_t* __restrict__ holes,
                                                       const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, const double* __restrict__ ox,
                                                       const double* __restrict__ oy, const double* __restrict__ oz,
                                                       long long m, double c, double* __restrict__ sx,
                                                       double* __restrict__ sy, double* __restrict__ sz,
                                                       uint8_t* __restrict__ visible) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (long long)gridDim.x * blockDim.x) {
        const int32_t h = holes[k];
        double hx = x[h], hy = y[h], hz = z[h];
        bool vis = true;
        if (r.hist) vis = source_pos(r, k, ox[j], oy[j], oz[j], c, hx, hy, hz);
        sx[j] = hx;
        sy[j] = hy;
        sz[j] = hz;
        visible[j] = vis;
    }
}

// The IDs of the alive holes (ids[holes[k]], or the state index without ids).
__global__ void k_gather_hole_ids(const int32_t* __restrict__ holes, long long nh, const uint64_t* __restrict__ ids,
                                  uint64_t* __restrict__ out) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nh; k += (long long)gridDim.x * blockDim.x)
        out[k] = ids ? ids[holes[k]] : (uint64_t)holes[k];
}

}  // namespace gx
