// gx_walk64.hip -- FP64 kernels: the reference's per-particle walk, the tree
// potential walk and the direct-sum oracle kernel.
//
// Compiled with -ffp-contract=off; the walk repeats internal/sim/bh/bh.go:106-155
// operation for operation, so its results are bit-identical to the reference's
// CPU walk on the same tree.  It is the "FP64 GPU walk" baseline: one thread per
// star in Morton order, stackless skip-pointer traversal.
#include "gx_common.h"
#include "gx_kernels.h"
#include "gx_pair.h"

namespace gx {

__global__ __launch_bounds__(256) void k_walk64(const double4* __restrict__ spos, const TravNode* __restrict__ trav,
                                                const Mult64* __restrict__ m64, const int32_t* __restrict__ order,
                                                Meta* meta, Batch batch, int quad, double* __restrict__ ax,
                                                double* __restrict__ ay, double* __restrict__ az) {
    const long long nt = batch.tOff[batch.B];
    if (meta->flags & kFlagTreeIncomplete) return;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long npp = 0, npc = 0;
    if (t < nt) {
        const long long k = target_star(batch, t);
        const int b = upper_index(batch.tOff, batch.B, t);
        const double eps2 = batch.eps2[b];
        const double4 p = spos[k];
        const double x = p.x, y = p.y, z = p.z;
        double sx = 0, sy = 0, sz = 0;
        for (int32_t i = batch.root[b]; i >= 0;) {
            const TravNode nd = trav[i];
            double dx = nd.x - x, dy = nd.y - y, dz = nd.z - z;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 >= nd.crit2) {
                const Mult64 mm = m64[i];
                double inv = 1 / (d2 + eps2);
                double inv3 = inv * sqrt(inv);
                double f = mm.m * inv3;
                if (quad) {
                    double qx = mm.qxx * dx + mm.qxy * dy + mm.qxz * dz;
                    double qy = mm.qxy * dx + mm.qyy * dy + mm.qyz * dz;
                    double qz = mm.qxz * dx + mm.qyz * dy + mm.qzz * dz;
                    double inv5 = inv3 * inv;
                    f += 2.5 * (dx * qx + dy * qy + dz * qz) * inv5 * inv;
                    sx -= qx * inv5;
                    sy -= qy * inv5;
                    sz -= qz * inv5;
                }
                sx += f * dx;
                sy += f * dy;
                sz += f * dz;
                npc++;
                i = nd.next;
                continue;
            }
            if (nd.first >= 0) {
                i = nd.first;
                continue;
            }
            for (int32_t j = nd.lo; j < nd.hi; j++) {
                if (j == (int32_t)k) continue;
                const double4 q = spos[j];
                double ddx = q.x - x, ddy = q.y - y, ddz = q.z - z;
                double r2 = ddx * ddx + ddy * ddy + ddz * ddz + eps2;
                double f = q.w / (r2 * sqrt(r2));
                sx += f * ddx;
                sy += f * ddy;
                sz += f * ddz;
            }
            npp += (unsigned long long)(nd.hi - nd.lo);
            i = nd.next;
        }
        const int32_t s = order[k];
        ax[s] = sx;
        ay[s] = sy;
        az[s] = sz;
    }
    npp = wave_sum_u64(npp);
    npc = wave_sum_u64(npc);
    if (lane_id() == 0 && (npp | npc)) {
        atomicAdd(&meta->pp, npp);
        atomicAdd(&meta->pc, npc);
    }
}

// Tree potential: walkPotential, internal/sim/bh/potential.go:45-81, minus
// the star's pairs with the holes (holePotential, :84-96).  Writes
// sm_k * psi_k - holes_k per star into partial[k] (sorted order); the
// host-side library sums them in a fixed order.
__global__ __launch_bounds__(256) void k_potential64(const double4* __restrict__ spos,
                                                     const TravNode* __restrict__ trav,
                                                     const Mult64* __restrict__ m64, Meta* meta, Batch batch,
                                                     int quad, const int32_t* __restrict__ order,
                                                     const int32_t* __restrict__ holes,
                                                     const double* __restrict__ px, const double* __restrict__ py,
                                                     const double* __restrict__ pz, const double* __restrict__ mass,
                                                     const uint8_t* __restrict__ kind, int dmOff, double c,
                                                     double* __restrict__ partial) {
    const long long nst = meta->nstars;
    if (meta->flags & kFlagTreeIncomplete) return;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nst) return;
    const int b = upper_index(batch.starOff, batch.B, k);
    const double eps2 = batch.eps2[b];
    const double4 p = spos[k];
    const double x = p.x, y = p.y, z = p.z;
    double psi = 0;
    for (int32_t i = batch.root[b]; i >= 0;) {
        const TravNode nd = trav[i];
        double dx = nd.x - x, dy = nd.y - y, dz = nd.z - z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 >= nd.crit2) {
            const Mult64 mm = m64[i];
            double inv = 1 / (d2 + eps2);
            psi += mm.m * sqrt(inv);
            if (quad) {
                double dqd = mm.qxx * dx * dx + mm.qyy * dy * dy + mm.qzz * dz * dz +
                             2 * (mm.qxy * dx * dy + mm.qxz * dx * dz + mm.qyz * dy * dz);
                psi += 0.5 * dqd * inv * inv * sqrt(inv);
            }
            i = nd.next;
            continue;
        }
        if (nd.first >= 0) {
            i = nd.first;
            continue;
        }
        for (int32_t j = nd.lo; j < nd.hi; j++) {
            if (j == (int32_t)k) continue;
            const double4 q = spos[j];
            double ddx = q.x - x, ddy = q.y - y, ddz = q.z - z;
            psi += q.w / sqrt(ddx * ddx + ddy * ddy + ddz * ddz + eps2);
        }
        i = nd.next;
    }
    // pairs with the holes: the star's source mass (SourceMass, sim.go:93-98)
    // against each hole, PW radius of the hole (PairRS, pair.go:13-22)
    double hp = 0;
    const long long nh = meta->nholes;
    if (nh > 0) {
        const int32_t si = order[k];
        const double xi = px[si], yi = py[si], zi = pz[si];
        const double mi = (dmOff && kind[si] == kDark) ? 0.0 : mass[si];
        for (long long h = 0; h < nh; h++) {
            const int32_t hi = holes[h];
            const double dx = px[hi] - xi, dy = py[hi] - yi, dz = pz[hi] - zi;
            hp += pair_potential(mi, mass[hi], dx * dx + dy * dy + dz * dz, eps2, schwarzschild(mass[hi], c));
        }
    }
    partial[k] = p.w * psi - hp;
}

// The holes' side of potential.go:29-38: per hole, its pairs with every star
// (one workgroup per hole, tree sum in a fixed order) and with the other
// holes (sequential).  out[2h] = sum over stars, out[2h+1] = over holes.
__global__ __launch_bounds__(256) void k_hole_potential(const double4* __restrict__ spos, const Meta* __restrict__ meta,
                                                        const int32_t* __restrict__ holes,
                                                        const double* __restrict__ px, const double* __restrict__ py,
                                                        const double* __restrict__ pz, const double* __restrict__ mass,
                                                        double eps2, double c, double* __restrict__ out) {
    __shared__ double red[256];
    const long long h = blockIdx.x;
    const long long nh = meta->nholes, nst = meta->nstars;
    if (h >= nh) return;
    const int32_t i = holes[h];
    const double x = px[i], y = py[i], z = pz[i], m = mass[i];
    const double rs = schwarzschild(m, c);
    double acc = 0;
    for (long long k = threadIdx.x; k < nst; k += blockDim.x) {
        const double4 q = spos[k];
        const double dx = q.x - x, dy = q.y - y, dz = q.z - z;
        acc += pair_potential(m, q.w, dx * dx + dy * dy + dz * dz, eps2, rs);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double hh = 0;  // two holes: PairRS = 0
        for (long long g = 0; g < nh; g++) {
            const int32_t j = holes[g];
            if (j == i) continue;
            const double dx = px[j] - x, dy = py[j] - y, dz = pz[j] - z;
            hh += pair_potential(m, mass[j], dx * dx + dy * dy + dz * dz, eps2, 0);
        }
        out[2 * h] = red[0];
        out[2 * h + 1] = hh;
    }
}

// FP64 direct sum for a sample of targets (direct.go:25-48): every workgroup
// owns 256 targets and one slice of the sources; slices are summed in a fixed
// order by k_direct_reduce, so the result is deterministic.
constexpr int kDirTile = 256;

__global__ __launch_bounds__(256) void k_direct_sample(const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, const double* __restrict__ m,
                                                       const uint8_t* __restrict__ kind, long long n,
                                                       const long long* __restrict__ targets, int nt, double eps2,
                                                       int dmOff, int splits, double* __restrict__ partial) {
    __shared__ double4 src[kDirTile];
    const int tid = threadIdx.x;
    const int t = blockIdx.x * 256 + tid;
    const int split = blockIdx.y;
    const long long per = (n + splits - 1) / splits;
    const long long jlo = (long long)split * per;
    const long long jhi = min(n, jlo + per);
    long long ti = t < nt ? targets[t] : -1;
    if (ti >= n) ti = -1;  // out-of-range targets give zeros instead of a fault
    double xi = 0, yi = 0, zi = 0;
    bool hole = true;
    if (ti >= 0) {
        xi = x[ti];
        yi = y[ti];
        zi = z[ti];
        hole = kind[ti] == kHole;
    }
    double sx = 0, sy = 0, sz = 0;
    for (long long j0 = jlo; j0 < jhi; j0 += kDirTile) {
        long long j = j0 + tid;
        double4 s = make_double4(0, 0, 0, 0);
        if (j < jhi) {
            uint8_t kj = kind[j];
            bool skip = kj == kHole || (dmOff && kj == kDark);
            s = make_double4(x[j], y[j], z[j], skip ? -1.0 : m[j]);
        }
        __syncthreads();
        src[tid] = s;
        __syncthreads();
        int cnt = (int)min((long long)kDirTile, jhi - j0);
        for (int u = 0; u < cnt; u++) {
            double4 q = src[u];
            if (q.w < 0 || j0 + u == ti) continue;
            double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
            double re2 = dx * dx + dy * dy + dz * dz + eps2;
            double f = q.w / (re2 * sqrt(re2));
            sx += f * dx;
            sy += f * dy;
            sz += f * dz;
        }
    }
    if (t < nt) {
        if (hole) sx = sy = sz = 0;
        size_t o = ((size_t)split * nt + t) * 3;
        partial[o] = sx;
        partial[o + 1] = sy;
        partial[o + 2] = sz;
    }
}

// Exact pair-model potential, PotentialEnergyParallel (diag.go:44-64): every
// pair i < j of the state, alive or not, with source masses (dark matter off
// pulls nothing) and PairRS from the raw masses (Paczynski-Wiita star-hole
// pairs for c > 0).  A workgroup owns 256 rows i and one slice of the columns
// j; partial[slice][i] holds row i's sum over the slice, in column order.
__global__ __launch_bounds__(256) void k_potential_exact(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z, const double* __restrict__ m,
                                                         const uint8_t* __restrict__ kind, long long n, double eps2,
                                                         int dmOff, double c, int slices,
                                                         double* __restrict__ partial) {
    __shared__ double4 src[kDirTile];  // x, y, z, source mass
    __shared__ double raw[kDirTile];   // mass (for the hole's radius)
    __shared__ int knd[kDirTile];
    const int tid = threadIdx.x;
    const long long i = (long long)blockIdx.x * 256 + tid;
    const long long per = (n + slices - 1) / slices;
    const long long jlo = (long long)blockIdx.y * per;
    const long long jhi = min(n, jlo + per);
    double xi = 0, yi = 0, zi = 0, mi = 0, rmi = 0;
    int ki = 0;
    if (i < n) {
        xi = x[i];
        yi = y[i];
        zi = z[i];
        ki = kind[i];
        rmi = m[i];
        mi = (dmOff && ki == kDark) ? 0.0 : rmi;
    }
    double part = 0;
    // columns j > i only: the workgroup's first row bounds where its tiles start
    const long long first = max(jlo, (long long)blockIdx.x * 256 + 1);
    for (long long j0 = first; j0 < jhi; j0 += kDirTile) {
        const long long j = j0 + tid;
        __syncthreads();
        if (j < jhi) {
            const int kj = kind[j];
            const double mj = m[j];
            src[tid] = make_double4(x[j], y[j], z[j], (dmOff && kj == kDark) ? 0.0 : mj);
            raw[tid] = mj;
            knd[tid] = kj;
        }
        __syncthreads();
        const int cnt = (int)min((long long)kDirTile, jhi - j0);
        if (i < n)
            for (int u = 0; u < cnt; u++) {
                if (j0 + u <= i) continue;
                const double4 q = src[u];
                const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const double rs = pair_rs(ki, rmi, knd[u], raw[u], c);
                part += pair_potential(mi, q.w, dx * dx + dy * dy + dz * dz, eps2, rs);
            }
    }
    if (i < n) partial[(size_t)blockIdx.y * n + i] = part;
}

__global__ void k_direct_reduce(const double* __restrict__ partial, int nt, int splits, double* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double a[3] = {0, 0, 0};
    for (int s = 0; s < splits; s++)
        for (int c = 0; c < 3; c++) a[c] += partial[((size_t)s * nt + t) * 3 + c];
    for (int c = 0; c < 3; c++) out[(size_t)t * 3 + c] = a[c];
}

}  // namespace gx
