/*
 * bh_oracle.c -- CPU restatement of the gravitas Barnes-Hut force step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP backend is
 * checked against (tests/, the smoke check and bench.py's cpu_baseline leg).
 * No product code path links, loads or calls it.
 *
 * It restates, step by step and in the same floating-point order, the Go
 * reference (paths relative to the gravitas repository):
 *   collect              internal/sim/bh/build.go:62-113
 *   keys                 internal/sim/bh/build.go:19-31, morton.go:7-21
 *   radixSort            internal/sim/bh/morton.go:25-49
 *   sorted copies        internal/sim/bh/build.go:34-41, internal/sim/sim.go:93-98
 *   buildNode/split      internal/sim/bh/build.go:134-172, lowerBound :340-350,
 *                        offset :353-358
 *   finish/quadrupole    internal/sim/bh/build.go:176-217, :326-337
 *   walk                 internal/sim/bh/bh.go:106-155
 *   Accelerations        internal/sim/bh/bh.go:76-102 (CPU branch)
 *   walkPotential        internal/sim/bh/potential.go:45-81, potentialWith :16-42
 *   direct.accRange      internal/sim/direct/direct.go:25-48
 *   ParallelFor          internal/sim/parallel.go:15-40
 *   kick / drift         internal/sim/sim.go:302-343 (Newtonian branch, c <= 0)
 *   PotentialEnergyParallel, kinetic, Momentum
 *                        internal/sim/diag.go:44-64, relativity.go:24-26, diag.go:100-109
 *   PairPotential        internal/sim/pair.go:37-47, SchwarzschildRadius
 *                        internal/sim/relativity.go:49-54
 *   stepSplit            internal/sim/split.go:27-199 (classify, holePull,
 *                        outerHoleForces, fastForces, kicks and drifts),
 *                        PairForce/PairRS internal/sim/pair.go:13-34
 *   frame dragging       internal/sim/framedrag.go:20-143
 *   retarded gravity     internal/sim/retarded.go:62-209 (holeHistory,
 *                        sourcePos, holeAppeared, recordHoles)
 *   rungs, block steps   internal/sim/blockstep.go:23-90
 *
 * Build with -O2 -ffp-contract=off (oracle/Makefile): Go's gc compiler on
 * amd64 does not fuse a*b+c, so the same operations in the same order give
 * the same IEEE-754 doubles.  The reference builds its tree serially below
 * 2^15 stars and in parallel above; both give the same tree and bit-identical
 * forces (internal/sim/bh/bh_test.go:378-394), so the serial recursive build
 * here stands for both.  The walk runs on pthreads with ParallelFor's dynamic
 * chunking; its result is independent of the thread count.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MORTON_BITS 21
#define LEAF_SIZE 24
#define DEFAULT_THETA 0.9
#define DIAG_THETA 0.9
#define GRAIN 128

enum { KIND_STAR = 0, KIND_HOLE = 1, KIND_DARK = 2, KIND_GAS = 3 };

typedef struct {
    double x, y, z, m;
    double qxx, qyy, qzz, qxy, qxz, qyz;
    double crit2;
    int32_t first, next, lo, hi;
} or_node;

typedef struct {
    int64_t n_stars, n_nodes, cap_nodes, cap_stars, n_holes, cap_holes;
    or_node *nodes;
    uint64_t *keys, *tmpk;
    int32_t *order, *tmpi, *holes;
    double *sx, *sy, *sz, *sm;
    double theta;
    int monopole;
    int workers;
} or_solver;

/* Go's math.Min / math.Max: NaN propagates, -0 < +0. */
static double go_min(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && b == 0) return signbit(a) ? a : b;
    return a < b ? a : b;
}
static double go_max(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && b == 0) return signbit(a) ? b : a;
    return a > b ? a : b;
}

/* ------------------------------------------------------------------ */
/* Morton keys: morton.go:7-21                                          */
/* ------------------------------------------------------------------ */
uint64_t or_spread(uint64_t v) {
    v &= (1ull << MORTON_BITS) - 1;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

uint64_t or_morton_key(uint64_t x, uint64_t y, uint64_t z) {
    return or_spread(x) << 2 | or_spread(y) << 1 | or_spread(z);
}

/* cell(): build.go:23-25.  Go's builtin min on floats propagates NaN. */
static uint64_t cell(double v, double lo, double scale) {
    double t = go_max(0, (v - lo) * scale);
    double c = isnan(t) ? t : (t < 2097151.0 ? t : 2097151.0);
    return (uint64_t)c;
}

/* ------------------------------------------------------------------ */
/* Stable LSD radix sort, 8-bit digits: morton.go:25-49                 */
/* ------------------------------------------------------------------ */
void or_radix_sort(uint64_t *keys, int32_t *idx, uint64_t *tmpk, int32_t *tmpi, int64_t n) {
    if (n <= 0) return;
    int64_t count[256];
    for (unsigned shift = 0; shift < 64; shift += 8) {
        memset(count, 0, sizeof count);
        for (int64_t i = 0; i < n; i++) count[(keys[i] >> shift) & 0xff]++;
        if (count[(keys[0] >> shift) & 0xff] == n) continue; /* all keys share this digit */
        int64_t sum = 0;
        for (int d = 0; d < 256; d++) {
            int64_t c = count[d];
            count[d] = sum;
            sum += c;
        }
        for (int64_t i = 0; i < n; i++) {
            uint64_t d = (keys[i] >> shift) & 0xff;
            tmpk[count[d]] = keys[i];
            tmpi[count[d]] = idx[i];
            count[d]++;
        }
        memcpy(keys, tmpk, (size_t)n * sizeof *keys);
        memcpy(idx, tmpi, (size_t)n * sizeof *idx);
    }
}

/* ------------------------------------------------------------------ */
/* Solver buffers                                                       */
/* ------------------------------------------------------------------ */
or_solver *or_create(void) {
    or_solver *s = calloc(1, sizeof *s);
    s->workers = 1;
    return s;
}

void or_destroy(or_solver *s) {
    if (!s) return;
    free(s->nodes); free(s->keys); free(s->tmpk); free(s->order); free(s->tmpi);
    free(s->holes); free(s->sx); free(s->sy); free(s->sz); free(s->sm);
    free(s);
}

static void reserve_stars(or_solver *s, int64_t n) {
    if (n <= s->cap_stars) return;
    int64_t c = n + n / 4 + 16;
    s->keys = realloc(s->keys, c * sizeof *s->keys);
    s->tmpk = realloc(s->tmpk, c * sizeof *s->tmpk);
    s->order = realloc(s->order, c * sizeof *s->order);
    s->tmpi = realloc(s->tmpi, c * sizeof *s->tmpi);
    s->sx = realloc(s->sx, c * sizeof *s->sx);
    s->sy = realloc(s->sy, c * sizeof *s->sy);
    s->sz = realloc(s->sz, c * sizeof *s->sz);
    s->sm = realloc(s->sm, c * sizeof *s->sm);
    s->cap_stars = c;
}

static int32_t new_node(or_solver *s) {
    if (s->n_nodes == s->cap_nodes) {
        s->cap_nodes = s->cap_nodes ? s->cap_nodes * 2 : 1024;
        s->nodes = realloc(s->nodes, s->cap_nodes * sizeof *s->nodes);
    }
    memset(&s->nodes[s->n_nodes], 0, sizeof(or_node));
    return (int32_t)s->n_nodes++;
}

/* ------------------------------------------------------------------ */
/* Tree: build.go:134-217, 326-358                                      */
/* ------------------------------------------------------------------ */
static int32_t lower_bound(const uint64_t *keys, int32_t lo, int32_t hi, unsigned shift, uint64_t oct) {
    while (lo < hi) {
        int32_t mid = (int32_t)((uint32_t)(lo + hi) >> 1);
        if (((keys[mid] >> shift) & 7) < oct)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

static int32_t split(const or_solver *s, int32_t lo, int32_t hi, int level, int32_t bounds[9]) {
    unsigned shift = (unsigned)(3 * (MORTON_BITS - 1 - level));
    bounds[0] = lo;
    bounds[8] = hi;
    for (uint64_t oct = 1; oct < 8; oct++) bounds[oct] = lower_bound(s->keys, bounds[oct - 1], hi, shift, oct);
    int32_t children = 0;
    for (int oct = 0; oct < 8; oct++)
        if (bounds[oct + 1] > bounds[oct]) children++;
    return children;
}

static double offset(int oct, int bit) { return ((oct >> bit) & 1) ? 1.0 : -1.0; }

static void quad_add(double q[6], double m, double dx, double dy, double dz) {
    double r2 = dx * dx + dy * dy + dz * dz;
    q[0] += m * (3 * dx * dx - r2);
    q[1] += m * (3 * dy * dy - r2);
    q[2] += m * (3 * dz * dz - r2);
    q[3] += m * 3 * dx * dy;
    q[4] += m * 3 * dx * dz;
    q[5] += m * 3 * dy * dz;
}

static or_node finish(const or_solver *s, int32_t lo, int32_t hi, int32_t first, int32_t children,
                      double cx, double cy, double cz, double half, int32_t next, double theta) {
    double x = 0, y = 0, z = 0, m = 0;
    if (first < 0) {
        for (int32_t j = lo; j < hi; j++) {
            double mj = s->sm[j];
            x = x + mj * s->sx[j];
            y = y + mj * s->sy[j];
            z = z + mj * s->sz[j];
            m = m + mj;
        }
    } else {
        for (int32_t c = first; c < first + children; c++) {
            const or_node *ch = &s->nodes[c];
            x = x + ch->m * ch->x;
            y = y + ch->m * ch->y;
            z = z + ch->m * ch->z;
            m = m + ch->m;
        }
    }
    if (m > 0) {
        x = x / m; y = y / m; z = z / m;
    } else {
        x = cx; y = cy; z = cz;
    }
    double q[6] = {0, 0, 0, 0, 0, 0};
    if (first < 0) {
        for (int32_t j = lo; j < hi; j++) quad_add(q, s->sm[j], s->sx[j] - x, s->sy[j] - y, s->sz[j] - z);
    } else {
        for (int32_t c = first; c < first + children; c++) {
            const or_node *ch = &s->nodes[c];
            q[0] = q[0] + ch->qxx; q[1] = q[1] + ch->qyy; q[2] = q[2] + ch->qzz;
            q[3] = q[3] + ch->qxy; q[4] = q[4] + ch->qxz; q[5] = q[5] + ch->qyz;
            quad_add(q, ch->m, ch->x - x, ch->y - y, ch->z - z);
        }
    }
    double delta = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    double crit = 2 * half / theta + delta;
    or_node nd;
    nd.x = x; nd.y = y; nd.z = z; nd.m = m;
    nd.qxx = q[0]; nd.qyy = q[1]; nd.qzz = q[2]; nd.qxy = q[3]; nd.qxz = q[4]; nd.qyz = q[5];
    nd.crit2 = crit * crit;
    nd.first = first; nd.next = next; nd.lo = lo; nd.hi = hi;
    return nd;
}

static void build_node(or_solver *s, int32_t ni, int32_t lo, int32_t hi, int level,
                       double cx, double cy, double cz, double half, int32_t next) {
    int32_t first = -1, children = 0;
    if (hi - lo > LEAF_SIZE && level < MORTON_BITS) {
        int32_t bounds[9];
        children = split(s, lo, hi, level, bounds);
        first = (int32_t)s->n_nodes;
        for (int32_t c = 0; c < children; c++) new_node(s);
        int32_t c = first;
        double q = half / 2;
        for (int oct = 0; oct < 8; oct++) {
            int32_t clo = bounds[oct], chi = bounds[oct + 1];
            if (chi == clo) continue;
            int32_t child_next = next;
            if (c < first + children - 1) child_next = c + 1;
            build_node(s, c, clo, chi, level + 1, cx + offset(oct, 2) * q, cy + offset(oct, 1) * q,
                       cz + offset(oct, 0) * q, q, child_next);
            c++;
        }
    }
    s->nodes[ni] = finish(s, lo, hi, first, children, cx, cy, cz, half, next, s->theta);
}

/* ------------------------------------------------------------------ */
/* ParallelFor: parallel.go:15-40 (dynamic chunks, 8 per worker)        */
/* ------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int64_t lo, int64_t hi);
typedef struct {
    range_fn fn;
    void *ctx;
    int64_t n, chunk, chunks;
    atomic_llong next;
} pf_job;

static void *pf_worker(void *arg) {
    pf_job *j = arg;
    for (;;) {
        int64_t c = atomic_fetch_add(&j->next, 1);
        if (c >= j->chunks) return NULL;
        int64_t lo = c * j->chunk, hi = lo + j->chunk < j->n ? lo + j->chunk : j->n;
        j->fn(j->ctx, lo, hi);
    }
}

static void parallel_for(int64_t n, int64_t grain, int workers, range_fn fn, void *ctx) {
    if (workers <= 0) workers = 1;
    if (n <= grain || workers == 1) {
        fn(ctx, 0, n);
        return;
    }
    pf_job j;
    j.fn = fn; j.ctx = ctx; j.n = n;
    j.chunk = n / ((int64_t)workers * 8);
    if (j.chunk < grain) j.chunk = grain;
    j.chunks = (n + j.chunk - 1) / j.chunk;
    atomic_init(&j.next, 0);
    int nt = workers < j.chunks ? workers : (int)j.chunks;
    pthread_t th[256];
    if (nt > 256) nt = 256;
    for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, pf_worker, &j);
    pf_worker(&j);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------ */
/* build: build.go:11-51 with collect :62-113                           */
/* ------------------------------------------------------------------ */
/* Returns the number of tree stars; fills order/keys/sorted copies/nodes. */
int64_t or_build(or_solver *s, const double *px, const double *py, const double *pz, const double *mass,
                 const uint8_t *kind, const uint8_t *alive, int64_t n, double theta, int dm_off) {
    s->theta = theta;
    s->n_nodes = 0;
    /* collect: holes, stars in state order, star bounding box */
    s->n_holes = 0;
    for (int64_t i = 0; i < n; i++) {
        if (alive[i] && kind[i] == KIND_HOLE) {
            if (s->n_holes == s->cap_holes) {
                s->cap_holes = s->cap_holes ? 2 * s->cap_holes : 16;
                s->holes = realloc(s->holes, s->cap_holes * sizeof *s->holes);
            }
            s->holes[s->n_holes++] = (int32_t)i;
        }
    }
    reserve_stars(s, n);
    double minX = INFINITY, minY = INFINITY, minZ = INFINITY;
    double maxX = -INFINITY, maxY = -INFINITY, maxZ = -INFINITY;
    int64_t ns = 0;
    for (int64_t i = 0; i < n; i++) {
        if (!(alive[i] && kind[i] != KIND_HOLE)) continue;
        s->order[ns++] = (int32_t)i;
        minX = go_min(minX, px[i]); maxX = go_max(maxX, px[i]);
        minY = go_min(minY, py[i]); maxY = go_max(maxY, py[i]);
        minZ = go_min(minZ, pz[i]); maxZ = go_max(maxZ, pz[i]);
    }
    s->n_stars = ns;
    if (ns == 0) return 0;

    double side = go_max(maxX - minX, go_max(maxY - minY, maxZ - minZ));
    side = go_max(side * (1 + 1e-9), 1e-12);
    double scale = (double)(1 << MORTON_BITS) / side;
    for (int64_t k = 0; k < ns; k++) {
        int32_t i = s->order[k];
        s->keys[k] = or_morton_key(cell(px[i], minX, scale), cell(py[i], minY, scale), cell(pz[i], minZ, scale));
    }
    or_radix_sort(s->keys, s->order, s->tmpk, s->tmpi, ns);
    for (int64_t k = 0; k < ns; k++) {
        int32_t i = s->order[k];
        s->sx[k] = px[i];
        s->sy[k] = py[i];
        s->sz[k] = pz[i];
        s->sm[k] = (kind[i] == KIND_DARK && dm_off) ? 0.0 : mass[i];
    }
    double half = side / 2;
    new_node(s);
    build_node(s, 0, 0, (int32_t)ns, 0, minX + half, minY + half, minZ + half, half, -1);
    return ns;
}

/* ------------------------------------------------------------------ */
/* walk: bh.go:106-155                                                  */
/* ------------------------------------------------------------------ */
void or_walk(const or_solver *s, double x, double y, double z, int32_t self, double eps2, double out[3]) {
    double ax = 0, ay = 0, az = 0;
    if (s->n_nodes == 0) {
        out[0] = out[1] = out[2] = 0;
        return;
    }
    const or_node *nodes = s->nodes;
    const double *sx = s->sx, *sy = s->sy, *sz = s->sz, *sm = s->sm;
    int quad = !s->monopole;
    for (int32_t i = 0; i >= 0;) {
        const or_node *nd = &nodes[i];
        double dx = nd->x - x, dy = nd->y - y, dz = nd->z - z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 >= nd->crit2) {
            double inv = 1 / (d2 + eps2);
            double inv3 = inv * sqrt(inv);
            double f = nd->m * inv3;
            if (quad) {
                double qx = nd->qxx * dx + nd->qxy * dy + nd->qxz * dz;
                double qy = nd->qxy * dx + nd->qyy * dy + nd->qyz * dz;
                double qz = nd->qxz * dx + nd->qyz * dy + nd->qzz * dz;
                double inv5 = inv3 * inv;
                f += 2.5 * (dx * qx + dy * qy + dz * qz) * inv5 * inv;
                ax -= qx * inv5;
                ay -= qy * inv5;
                az -= qz * inv5;
            }
            ax += f * dx;
            ay += f * dy;
            az += f * dz;
            i = nd->next;
            continue;
        }
        if (nd->first >= 0) {
            i = nd->first;
            continue;
        }
        for (int32_t j = nd->lo; j < nd->hi; j++) {
            if (j == self) continue;
            double ddx = sx[j] - x, ddy = sy[j] - y, ddz = sz[j] - z;
            double r2 = ddx * ddx + ddy * ddy + ddz * ddz + eps2;
            double f = sm[j] / (r2 * sqrt(r2));
            ax += f * ddx;
            ay += f * ddy;
            az += f * ddz;
        }
        i = nd->next;
    }
    out[0] = ax; out[1] = ay; out[2] = az;
}

/* walkPotential: potential.go:45-81 */
double or_walk_potential(const or_solver *s, double x, double y, double z, int32_t self, double eps2) {
    if (s->n_nodes == 0) return 0;
    const or_node *nodes = s->nodes;
    int quad = !s->monopole;
    double psi = 0;
    for (int32_t i = 0; i >= 0;) {
        const or_node *nd = &nodes[i];
        double dx = nd->x - x, dy = nd->y - y, dz = nd->z - z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 >= nd->crit2) {
            double inv = 1 / (d2 + eps2);
            psi += nd->m * sqrt(inv);
            if (quad) {
                double dqd = nd->qxx * dx * dx + nd->qyy * dy * dy + nd->qzz * dz * dz +
                             2 * (nd->qxy * dx * dy + nd->qxz * dx * dz + nd->qyz * dy * dz);
                psi += 0.5 * dqd * inv * inv * sqrt(inv);
            }
            i = nd->next;
            continue;
        }
        if (nd->first >= 0) {
            i = nd->first;
            continue;
        }
        for (int32_t j = nd->lo; j < nd->hi; j++) {
            if (j == self) continue;
            double ddx = s->sx[j] - x, ddy = s->sy[j] - y, ddz = s->sz[j] - z;
            psi += s->sm[j] / sqrt(ddx * ddx + ddy * ddy + ddz * ddz + eps2);
        }
        i = nd->next;
    }
    return psi;
}

/* ------------------------------------------------------------------ */
/* Accelerations (CPU branch): bh.go:76-102                             */
/* ------------------------------------------------------------------ */
typedef struct {
    const or_solver *s;
    double eps2;
    double *ax, *ay, *az;
} acc_ctx;

static void acc_range(void *p, int64_t lo, int64_t hi) {
    acc_ctx *c = p;
    const or_solver *s = c->s;
    for (int64_t k = lo; k < hi; k++) {
        int32_t i = s->order[k];
        double a[3];
        or_walk(s, s->sx[k], s->sy[k], s->sz[k], (int32_t)k, c->eps2, a);
        c->ax[i] = a[0]; c->ay[i] = a[1]; c->az[i] = a[2];
    }
}

/* Accelerations of the last build into state-order arrays.  Holes get 0;
 * dead bodies are left untouched, as in the reference. */
void or_accelerations_built(or_solver *s, double eps, double *ax, double *ay, double *az) {
    acc_ctx c = {s, eps * eps, ax, ay, az};
    parallel_for(s->n_stars, GRAIN, s->workers, acc_range, &c);
    for (int64_t h = 0; h < s->n_holes; h++) {
        int32_t i = s->holes[h];
        ax[i] = ay[i] = az[i] = 0;
    }
}

/* Accelerations: theta == 0 means DefaultTheta (bh.go:69-74). */
void or_accelerations(or_solver *s, const double *px, const double *py, const double *pz, const double *mass,
                      const uint8_t *kind, const uint8_t *alive, int64_t n, double theta, double eps, int dm_off,
                      int monopole, int workers, double *ax, double *ay, double *az) {
    s->monopole = monopole;
    s->workers = workers;
    or_build(s, px, py, pz, mass, kind, alive, n, theta == 0 ? DEFAULT_THETA : theta, dm_off);
    or_accelerations_built(s, eps, ax, ay, az);
}

/* Accelerations of the sorted stars (sorted order), for walk-level checks. */
void or_accelerations_sorted(or_solver *s, double eps, double *ax, double *ay, double *az) {
    double eps2 = eps * eps;
    for (int64_t k = 0; k < s->n_stars; k++) {
        double a[3];
        or_walk(s, s->sx[k], s->sy[k], s->sz[k], (int32_t)k, eps2, a);
        ax[k] = a[0]; ay[k] = a[1]; az[k] = a[2];
    }
}

/* PairPotential / SchwarzschildRadius: pair.go:37-47, relativity.go:49-54 */
static double pair_potential(double mi, double mj, double r2, double eps2, double rs) {
    double re = sqrt(r2 + eps2);
    if (rs == 0) return -mi * mj / re;
    if (re >= 1.5 * rs) return -mi * mj / (re - rs);
    return mi * mj * (-2 / rs + 4 / (rs * rs) * (re - 1.5 * rs));
}
static double schwarzschild(double m, double c) { return c <= 0 ? 0 : 2 * m / (c * c); }

/* PotentialEnergy: potential.go:12-42, theta capped at diagTheta.  The
 * per-star psi values are summed in sorted order (the reference sums chunk
 * results in completion order, so only round-off may differ). */
double or_potential_energy(or_solver *s, const double *px, const double *py, const double *pz, const double *mass,
                           const uint8_t *kind, const uint8_t *alive, int64_t n, double theta, double eps, int dm_off,
                           int monopole, double c) {
    s->monopole = monopole;
    double th = theta == 0 ? DEFAULT_THETA : theta;
    if (th > DIAG_THETA) th = DIAG_THETA;
    or_build(s, px, py, pz, mass, kind, alive, n, th, dm_off);
    double eps2 = eps * eps, sum = 0;
    for (int64_t k = 0; k < s->n_stars; k++) {
        int32_t i = s->order[k];
        double hp = 0;
        for (int64_t h = 0; h < s->n_holes; h++) {
            int32_t hi = s->holes[h];
            if (hi == i) continue;
            double dx = px[hi] - px[i], dy = py[hi] - py[i], dz = pz[hi] - pz[i];
            double rs = (c <= 0) ? 0 : schwarzschild(mass[hi], c); /* PairRS: star-hole pair */
            double mi = (kind[i] == KIND_DARK && dm_off) ? 0 : mass[i];
            hp += pair_potential(mi, mass[hi], dx * dx + dy * dy + dz * dz, eps2, rs);
        }
        sum += s->sm[k] * or_walk_potential(s, s->sx[k], s->sy[k], s->sz[k], (int32_t)k, eps2) - hp;
    }
    for (int64_t h = 0; h < s->n_holes; h++) {
        int32_t i = s->holes[h];
        double rs = schwarzschild(mass[i], c);
        for (int64_t k = 0; k < s->n_stars; k++) {
            double dx = s->sx[k] - px[i], dy = s->sy[k] - py[i], dz = s->sz[k] - pz[i];
            sum -= pair_potential(mass[i], s->sm[k], dx * dx + dy * dy + dz * dz, eps2, rs);
        }
        double hp = 0; /* hole-hole pairs: PairRS gives 0 for two holes */
        for (int64_t g = 0; g < s->n_holes; g++) {
            int32_t j = s->holes[g];
            if (j == i) continue;
            double dx = px[j] - px[i], dy = py[j] - py[i], dz = pz[j] - pz[i];
            hp += pair_potential(mass[i], mass[j], dx * dx + dy * dy + dz * dz, eps2, 0);
        }
        sum -= hp;
    }
    return -sum / 2;
}

/* ------------------------------------------------------------------ */
/* Direct sum: direct.go:25-48                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *px, *py, *pz, *m;
    const uint8_t *kind;
    const int64_t *targets;
    int64_t n;
    double eps2;
    int dm_off;
    double *ax, *ay, *az;
} dir_ctx;

static void direct_one(const dir_ctx *c, int64_t i, double out[3]) {
    double xi = c->px[i], yi = c->py[i], zi = c->pz[i];
    double ax = 0, ay = 0, az = 0;
    if (c->kind[i] == KIND_HOLE) {
        out[0] = out[1] = out[2] = 0;
        return;
    }
    for (int64_t j = 0; j < c->n; j++) {
        if (j == i || c->kind[j] == KIND_HOLE || (c->dm_off && c->kind[j] == KIND_DARK)) continue;
        double dx = c->px[j] - xi, dy = c->py[j] - yi, dz = c->pz[j] - zi;
        double re2 = dx * dx + dy * dy + dz * dz + c->eps2;
        double f = c->m[j] / (re2 * sqrt(re2));
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
    }
    out[0] = ax; out[1] = ay; out[2] = az;
}

static void direct_range(void *p, int64_t lo, int64_t hi) {
    dir_ctx *c = p;
    for (int64_t t = lo; t < hi; t++) {
        int64_t i = c->targets ? c->targets[t] : t;
        double a[3];
        direct_one(c, i, a);
        c->ax[t] = a[0]; c->ay[t] = a[1]; c->az[t] = a[2];
    }
}

/* Direct-sum accelerations of targets[0..nt) (all bodies if targets is NULL,
 * then nt must equal n); out[t] belongs to targets[t]. */
void or_direct(const double *px, const double *py, const double *pz, const double *mass, const uint8_t *kind,
               int64_t n, const int64_t *targets, int64_t nt, double eps, int dm_off, int workers,
               double *ax, double *ay, double *az) {
    dir_ctx c = {px, py, pz, mass, kind, targets, n, eps * eps, dm_off, ax, ay, az};
    parallel_for(targets ? nt : n, 64, workers, direct_range, &c);
}

/* ------------------------------------------------------------------ */
/* Integrator pieces (sim.go:302-343, Newtonian) and diagnostics        */
/* ------------------------------------------------------------------ */
void or_kick(int64_t n, double *vx, double *vy, double *vz, const double *ax, const double *ay, const double *az,
             double h) {
    for (int64_t i = 0; i < n; i++) {
        vx[i] += h * ax[i];
        vy[i] += h * ay[i];
        vz[i] += h * az[i];
    }
}

void or_drift(int64_t n, double *px, double *py, double *pz, const double *vx, const double *vy, const double *vz,
              double dt) {
    for (int64_t i = 0; i < n; i++) {
        px[i] += dt * vx[i];
        py[i] += dt * vy[i];
        pz[i] += dt * vz[i];
    }
}

/* PotentialEnergyParallel, diag.go:44-64: the exact pair model over every
 * pair i < j of the state (alive or not, as the reference's loop), source
 * masses (dark matter off pulls nothing), PairRS from the raw masses.  The
 * reference adds chunk sums in completion order; here rows in order. */
static double pair_rs(int ki, double mi, int kj, double mj, double c);
double or_potential_exact(const double *px, const double *py, const double *pz, const double *mass,
                          const uint8_t *kind, int64_t n, double eps, int dm_off, double c) {
    double eps2 = eps * eps, sum = 0;
    for (int64_t i = 0; i < n; i++) {
        double mi = (kind[i] == KIND_DARK && dm_off) ? 0 : mass[i];
        double xi = px[i], yi = py[i], zi = pz[i];
        double part = 0;
        for (int64_t j = i + 1; j < n; j++) {
            double mj = (kind[j] == KIND_DARK && dm_off) ? 0 : mass[j];
            double dx = px[j] - xi, dy = py[j] - yi, dz = pz[j] - zi;
            double rs = pair_rs(kind[i], mass[i], kind[j], mass[j], c);
            part += pair_potential(mi, mj, dx * dx + dy * dy + dz * dz, eps2, rs);
        }
        sum += part;
    }
    return sum;
}

/* ------------------------------------------------------------------ */
/* Block time steps (gravity use of the gas rung scheme)                */
/* ------------------------------------------------------------------ */
/* rungFor, internal/sim/blockstep.go:37-45: the smallest r with T/2^r <= dt,
 * capped at max_rung. */
int or_rung_for(double dt, double T, int max_rung) {
    int r = 0;
    double x = T;
    while (x > dt && r < max_rung) {
        x *= 0.5;
        r++;
    }
    return r;
}

/* Bodies whose step ends at tick `now` (a step of rung r spans full >> r
 * ticks and starts at a multiple of that, blockstep.go:81-90). */
void or_block_active(int64_t n, const int8_t *rung, int now, int full, uint8_t *active) {
    for (int64_t i = 0; i < n; i++) active[i] = now % (full >> rung[i]) == 0;
}

/* Kicks at tick `now` of a global step T: the closing kick of each body whose
 * step ends (now > 0; kickGas(i, 0.5*step), blockstep.go:187-189), then, if
 * now < full, a new rung from the wanted step sqrt(crit / |a|) aligned to now
 * (startStep, :81-90) and its opening kick.  now = 0 starts every body.
 * Holes and dead bodies stay on rung 0.  crit = 2 eta eps (GADGET-2's
 * criterion; the reference has no gravity step criterion). */
void or_block_kick(int64_t n, double *vx, double *vy, double *vz, const double *ax, const double *ay,
                   const double *az, const uint8_t *kind, const uint8_t *alive, int8_t *rung, int now, int full,
                   int max_rung, double T, double crit) {
    for (int64_t i = 0; i < n; i++) {
        int r = rung[i];
        if (now > 0 && now % (full >> r) != 0) continue;
        if (now > 0) {
            double h = 0.5 * (T / (double)(1 << r));
            vx[i] += h * ax[i];
            vy[i] += h * ay[i];
            vz[i] += h * az[i];
        }
        if (now >= full) continue;
        r = 0;
        if (alive[i] && kind[i] != KIND_HOLE) {
            double x = ax[i], y = ay[i], z = az[i];
            double a = sqrt(x * x + y * y + z * z);
            r = or_rung_for(sqrt(crit / a), T, max_rung);
        }
        while (now % (full >> r) != 0) r++;
        rung[i] = (int8_t)r;
        double h = 0.5 * (T / (double)(1 << r));
        vx[i] += h * ax[i];
        vy[i] += h * ay[i];
        vz[i] += h * az[i];
    }
}

/* Diagnostics, internal/sim/diag.go, summed in state order as there:
 * out[0] sourceKinetic (:68-77, kinetic with c <= 0: m u^2 / (1 + 1)),
 * out[1..3] Momentum (:102-109), out[4..6] AngularMomentum (:112-119, x cross
 * (m v)), out[7] total mass and out[8..10] sum m x (CenterOfMass :122-134). */
void or_diagnostics(int64_t n, const double *px, const double *py, const double *pz, const double *vx,
                    const double *vy, const double *vz, const double *mass, const uint8_t *kind, int dm_off,
                    double *out) {
    for (int c = 0; c < 11; c++) out[c] = 0;
    for (int64_t i = 0; i < n; i++) {
        double m = mass[i], ux = vx[i], uy = vy[i], uz = vz[i], x = px[i], y = py[i], z = pz[i];
        if (!(kind[i] == KIND_DARK && dm_off)) out[0] += m * (ux * ux + uy * uy + uz * uz) / (1.0 + 1.0);
        double qx = ux * m, qy = uy * m, qz = uz * m;
        out[1] += qx;
        out[2] += qy;
        out[3] += qz;
        out[4] += y * qz - z * qy;
        out[5] += z * qx - x * qz;
        out[6] += x * qy - y * qx;
        out[7] += m;
        out[8] += x * m;
        out[9] += y * m;
        out[10] += z * m;
    }
}

/* ------------------------------------------------------------------ */
/* Multiple time stepping near black holes: internal/sim/split.go,      */
/* with frame dragging of spinning holes (framedrag.go); no gas,         */
/* instant gravity                                                       */
/* ------------------------------------------------------------------ */
/* PairForce, pair.go:26-34 */
static double pair_force(double m, double r2, double eps2, double rs) {
    double re2 = r2 + eps2;
    double re = sqrt(re2);
    if (rs == 0) return m / (re2 * re);
    double den = go_max(re - rs, rs / 2);
    return m / (den * den * re);
}
/* gamma, relativity.go:14-19 */
static double gamma_of(double ux, double uy, double uz, double c) {
    if (c <= 0) return 1;
    return sqrt(1 + (ux * ux + uy * uy + uz * uz) / (c * c));
}
/* PairRS, pair.go:13-22 */
static double pair_rs(int ki, double mi, int kj, double mj, double c) {
    if (c <= 0 || (ki == KIND_HOLE) == (kj == KIND_HOLE)) return 0;
    return ki == KIND_HOLE ? schwarzschild(mi, c) : schwarzschild(mj, c);
}

/* Retarded gravity of black holes, internal/sim/retarded.go: a ring buffer
 * of each hole's positions and coordinate velocities at outer steps, keyed
 * by body ID; a body sees a hole where it was at the retarded time. */
#define HIST_LEN 8192     /* historyLen, retarded.go:50-52 */
#define RETARD_TOL 1e-7   /* retardTolerance, :54-59 */
#define RETARD_MAX_IT 64  /* retardMaxIterations */
enum { GRAVITY_INSTANT = 0, GRAVITY_RETARDED = 1, GRAVITY_RETARDED_V = 2 };

typedef struct {
    uint64_t id;
    double t0, appeared, dt;
    int n, head;
    double *x, *y, *z, *vx, *vy, *vz;
} or_hist;

typedef struct {
    or_hist **h;
    int count, cap;
} or_retard;

or_retard *or_retard_new(void) { return calloc(1, sizeof(or_retard)); }
static void hist_free(or_hist *h) {
    free(h->x);
    free(h);
}
void or_retard_free(or_retard *r) {
    if (!r) return;
    for (int i = 0; i < r->count; i++) hist_free(r->h[i]);
    free(r->h);
    free(r);
}
int or_retard_count(const or_retard *r) { return r->count; }

static or_hist *hist_find(const or_retard *r, uint64_t id) {
    for (int i = 0; i < r->count; i++)
        if (r->h[i]->id == id) return r->h[i];
    return NULL;
}

/* push, retarded.go:75-83 */
static void hist_push(or_hist *h, double t, double x, double y, double z, double vx, double vy, double vz) {
    h->head = (h->head + 1) % HIST_LEN;
    h->x[h->head] = x;
    h->y[h->head] = y;
    h->z[h->head] = z;
    h->vx[h->head] = vx;
    h->vy[h->head] = vy;
    h->vz[h->head] = vz;
    if (h->n < HIST_LEN) h->n++;
    h->t0 = t - (double)(h->n - 1) * h->dt;
}

/* at, retarded.go:85-103: the state interpolated at t (extrapolated before
 * the oldest entry); 0 if the hole had not appeared yet */
static int hist_at(const or_hist *h, double t, double o[6]) {
    if (t < h->appeared) {
        for (int i = 0; i < 6; i++) o[i] = 0;
        return 0;
    }
    int oldest = (h->head - h->n + 1 + HIST_LEN) % HIST_LEN;
    if (t <= h->t0 || h->n == 1) {
        double d = t - h->t0;
        o[0] = h->x[oldest] + d * h->vx[oldest];
        o[1] = h->y[oldest] + d * h->vy[oldest];
        o[2] = h->z[oldest] + d * h->vz[oldest];
        o[3] = h->vx[oldest];
        o[4] = h->vy[oldest];
        o[5] = h->vz[oldest];
        return 1;
    }
    double f = (t - h->t0) / h->dt;
    int k = (int)f;
    if (k > h->n - 2) k = h->n - 2;
    double w = f - (double)k;
    int a = (oldest + k) % HIST_LEN, b = (oldest + k + 1) % HIST_LEN;
    const double *arr[6] = {h->x, h->y, h->z, h->vx, h->vy, h->vz};
    for (int i = 0; i < 6; i++) o[i] = arr[i][a] + w * (arr[i][b] - arr[i][a]);
    return 1;
}

/* history, retarded.go:163-179: create with the current state as first entry */
static or_hist *hist_create(or_retard *r, uint64_t id, double t, double dt, double x, double y, double z,
                            double ux, double uy, double uz, double c) {
    or_hist *h = calloc(1, sizeof(or_hist));
    h->id = id;
    h->appeared = -INFINITY;
    h->dt = dt;
    h->x = malloc(6 * HIST_LEN * sizeof(double));
    h->y = h->x + HIST_LEN;
    h->z = h->y + HIST_LEN;
    h->vx = h->z + HIST_LEN;
    h->vy = h->vx + HIST_LEN;
    h->vz = h->vy + HIST_LEN;
    double g = c <= 0 ? 1 : sqrt(1 + (ux * ux + uy * uy + uz * uz) / (c * c)); /* CoordinateVelocity */
    double ig = 1 / g;
    hist_push(h, t, x, y, z, ux * ig, uy * ig, uz * ig);
    if (r->count == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 8;
        r->h = realloc(r->h, (size_t)r->cap * sizeof(or_hist *));
    }
    r->h[r->count++] = h;
    return h;
}

static uint64_t body_id(const uint64_t *ids, int64_t i) { return ids ? ids[i] : (uint64_t)i; }

/* holeAppeared, retarded.go:157-161 (AddBody of a hole, sim.go:236-243) */
void or_hole_appeared(or_retard *r, const double *px, const double *py, const double *pz, const double *vx,
                      const double *vy, const double *vz, const uint64_t *ids, int64_t i, double time, double dt,
                      double c) {
    uint64_t id = body_id(ids, i);
    or_hist *h = hist_find(r, id);
    if (!h) h = hist_create(r, id, time, dt, px[i], py[i], pz[i], vx[i], vy[i], vz[i], c);
    h->appeared = time;
}

/* recordHoles, retarded.go:181-209, at the time after the step */
static void record_holes(or_retard *r, const double *px, const double *py, const double *pz, const double *vx,
                         const double *vy, const double *vz, const uint8_t *kind, const uint8_t *alive,
                         const uint64_t *ids, int64_t n, double time, double dt, double c, int mode) {
    if (!(mode != GRAVITY_INSTANT && c > 0)) {
        for (int i = 0; i < r->count; i++) hist_free(r->h[i]);
        r->count = 0;
        return;
    }
    int64_t nh = 0;
    for (int64_t i = 0; i < n; i++)
        if (alive[i] && kind[i] == KIND_HOLE) nh++;
    int seen = r->count != nh;
    for (int64_t i = 0; i < n; i++) {
        if (!(alive[i] && kind[i] == KIND_HOLE)) continue;
        uint64_t id = body_id(ids, i);
        or_hist *h = hist_find(r, id);
        if (h) {
            double g = c <= 0 ? 1 : sqrt(1 + (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) / (c * c));
            double ig = 1 / g;
            hist_push(h, time, px[i], py[i], pz[i], vx[i] * ig, vy[i] * ig, vz[i] * ig);
        } else {
            hist_create(r, id, time, dt, px[i], py[i], pz[i], vx[i], vy[i], vz[i], c);
        }
    }
    if (seen) { /* drop histories of holes that merged away */
        int keep = 0;
        for (int j = 0; j < r->count; j++) {
            int alive_hole = 0;
            for (int64_t i = 0; i < n && !alive_hole; i++)
                alive_hole = alive[i] && kind[i] == KIND_HOLE && body_id(ids, i) == r->h[j]->id;
            if (alive_hole)
                r->h[keep++] = r->h[j];
            else
                hist_free(r->h[j]);
        }
        r->count = keep;
    }
}

typedef struct {
    double *px, *py, *pz, *vx, *vy, *vz;
    const double *mass;
    const uint8_t *kind;
    int64_t n;
    double eps2, c;
    int dm_off;
    const int32_t *holes;
    int64_t nh;
    uint8_t *near;
    int32_t *nearList;
    int64_t nn;
    const uint8_t *alive;
    double *jx, *jy, *jz; /* hole spins (angular momentum), NULL: none */
    double *weights;      /* dragBack's per-hole weights, nh entries */
    or_hist **hists;      /* history of each hole slot (bindHistories), NULL: none */
    int mode;             /* Params.Gravity */
    double time, dt;      /* Simulation.Time during the step, Params.Dt */
} split_ctx;

static double source_mass(const split_ctx *q, int64_t i) {
    return (q->kind[i] == KIND_DARK && q->dm_off) ? 0.0 : q->mass[i];
}

/* holeTimescale, split.go:27-36 */
static double hole_timescale(const split_ctx *q, int64_t i, int64_t h) {
    double dx = q->px[i] - q->px[h], dy = q->py[i] - q->py[h], dz = q->pz[i] - q->pz[h];
    double r2 = dx * dx + dy * dy + dz * dz;
    double rs = pair_rs(q->kind[i], q->mass[i], KIND_HOLE, q->mass[h], q->c);
    double a = pair_force(q->mass[h], r2, q->eps2, rs) * sqrt(r2);
    if (a == 0) return INFINITY;
    return sqrt(sqrt(r2 + q->eps2) / a);
}

/* sourcePos, retarded.go:118-155: where a body at (x, y, z) sees hole slot
 * k now; 0 if its gravity has not arrived there yet */
static int source_pos(const split_ctx *q, int64_t k, double x, double y, double z, double o[3]) {
    int64_t h = q->holes[k];
    double hx = q->px[h], hy = q->py[h], hz = q->pz[h];
    o[0] = hx;
    o[1] = hy;
    o[2] = hz;
    if (!(q->mode != GRAVITY_INSTANT && q->c > 0)) return 1;
    const or_hist *hist = q->hists ? q->hists[k] : NULL;
    double c = q->c;
    double delay = sqrt((x - hx) * (x - hx) + (y - hy) * (y - hy) + (z - hz) * (z - hz)) / c;
    if (!hist || delay < q->dt) return 1;
    double st[6];
    for (int it = 0; it < RETARD_MAX_IT; it++) {
        if (!hist_at(hist, q->time - delay, st)) {
            o[0] = o[1] = o[2] = 0;
            return 0;
        }
        hx = st[0];
        hy = st[1];
        hz = st[2];
        double next = sqrt((x - hx) * (x - hx) + (y - hy) * (y - hy) + (z - hz) * (z - hz)) / c;
        int converged = fabs(next - delay) < RETARD_TOL;
        delay = next;
        if (converged) break;
    }
    hist_at(hist, q->time - delay, st);
    hx = st[0];
    hy = st[1];
    hz = st[2];
    if (q->mode == GRAVITY_RETARDED_V) {
        hx = hx + delay * st[3];
        hy = hy + delay * st[4];
        hz = hz + delay * st[5];
    }
    o[0] = hx;
    o[1] = hy;
    o[2] = hz;
    return 1;
}

/* holePull, split.go:120-137: the pull of every hole but i on body i, from
 * where the body sees it (sourcePos) */
static void hole_pull(const split_ctx *q, int64_t i, double out[3]) {
    double ax = 0, ay = 0, az = 0;
    for (int64_t k = 0; k < q->nh; k++) {
        int64_t h = q->holes[k];
        if (h == i) continue;
        double x = q->px[i], y = q->py[i], z = q->pz[i];
        double src[3];
        if (!source_pos(q, k, x, y, z, src)) continue;
        double dx = src[0] - x, dy = src[1] - y, dz = src[2] - z;
        double rs = pair_rs(q->kind[i], q->mass[i], KIND_HOLE, q->mass[h], q->c);
        double f = pair_force(q->mass[h], dx * dx + dy * dy + dz * dz, q->eps2, rs);
        ax = ax + f * dx;
        ay = ay + f * dy;
        az = az + f * dz;
    }
    out[0] = ax; out[1] = ay; out[2] = az;
}

/* classify, split.go:39-80: near flags, near list (index order), substeps */
static int classify(split_ctx *q, double dt) {
    double limit = 20.0 * dt, tau_min = INFINITY;
    q->nn = 0;
    for (int64_t i = 0; i < q->n; i++) {
        int near = 0;
        if (q->kind[i] != KIND_HOLE) {
            for (int64_t k = 0; k < q->nh; k++) {
                double tau = hole_timescale(q, i, q->holes[k]);
                if (tau < limit) {
                    near = 1;
                    tau_min = go_min(tau_min, tau);
                }
            }
        }
        q->near[i] = (uint8_t)near;
        if (near) q->nearList[q->nn++] = (int32_t)i;
    }
    for (int64_t a = 0; a < q->nh; a++)
        for (int64_t b = a + 1; b < q->nh; b++) {
            double tau = hole_timescale(q, q->holes[a], q->holes[b]);
            if (tau < limit) tau_min = go_min(tau_min, tau);
        }
    if (isinf(tau_min)) return 1;
    int k = 1;
    while ((double)k * tau_min < 20.0 * dt && k < 256) k *= 2;
    return k;
}

/* outerHoleForces, split.go:82-117 (hole sums in index order) */
static void outer_hole_forces(const split_ctx *q, double *fx, double *fy, double *fz) {
    for (int64_t i = 0; i < q->n; i++) {
        fx[i] = fy[i] = fz[i] = 0;
        if (q->kind[i] == KIND_HOLE || q->near[i]) continue;
        double a[3];
        hole_pull(q, i, a);
        fx[i] = a[0]; fy[i] = a[1]; fz[i] = a[2];
    }
    for (int64_t k = 0; k < q->nh; k++) {
        int64_t h = q->holes[k];
        double x = q->px[h], y = q->py[h], z = q->pz[h], rs = schwarzschild(q->mass[h], q->c);
        double ax = 0, ay = 0, az = 0;
        for (int64_t i = 0; i < q->n; i++) {
            if (q->kind[i] == KIND_HOLE || q->near[i]) continue;
            double dx = q->px[i] - x, dy = q->py[i] - y, dz = q->pz[i] - z;
            double f = pair_force(source_mass(q, i), dx * dx + dy * dy + dz * dz, q->eps2, rs);
            ax = ax + f * dx;
            ay = ay + f * dy;
            az = az + f * dz;
        }
        fx[h] = fx[h] + ax;
        fy[h] = fy[h] + ay;
        fz[h] = fz[h] + az;
    }
}

/* fastForces, split.go:141-160 */
static void fast_forces(const split_ctx *q, double *fx, double *fy, double *fz) {
    for (int64_t j = 0; j < q->nn; j++) {
        int64_t i = q->nearList[j];
        double a[3];
        hole_pull(q, i, a);
        fx[i] = a[0]; fy[i] = a[1]; fz[i] = a[2];
    }
    for (int64_t k = 0; k < q->nh; k++) {
        int64_t h = q->holes[k];
        double x = q->px[h], y = q->py[h], z = q->pz[h], rs = schwarzschild(q->mass[h], q->c);
        double a[3];
        hole_pull(q, h, a);
        double ax = a[0], ay = a[1], az = a[2];
        for (int64_t j = 0; j < q->nn; j++) {
            int64_t i = q->nearList[j];
            double dx = q->px[i] - x, dy = q->py[i] - y, dz = q->pz[i] - z;
            double f = pair_force(source_mass(q, i), dx * dx + dy * dy + dz * dz, q->eps2, rs);
            ax = ax + f * dx;
            ay = ay + f * dy;
            az = az + f * dz;
        }
        fx[h] = ax; fy[h] = ay; fz[h] = az;
    }
}

static void kick_idx(const split_ctx *q, const int32_t *idx, int64_t cnt, const double *fx, const double *fy,
                     const double *fz, double h) {
    for (int64_t j = 0; j < cnt; j++) {
        int64_t i = idx[j];
        q->vx[i] += h * fx[i];
        q->vy[i] += h * fy[i];
        q->vz[i] += h * fz[i];
    }
}

/* driftIdx, split.go:214-222 */
static void drift_idx(const split_ctx *q, const int32_t *idx, int64_t cnt, double dt) {
    for (int64_t j = 0; j < cnt; j++) {
        int64_t i = idx[j];
        double f = dt / gamma_of(q->vx[i], q->vy[i], q->vz[i], q->c);
        q->px[i] += f * q->vx[i];
        q->py[i] += f * q->vy[i];
        q->pz[i] += f * q->vz[i];
    }
}

/* spinning, framedrag.go:20-32: frame dragging acts if c > 0 and an alive
 * hole has a spin */
static int spinning(const split_ctx *q) {
    if (q->c <= 0 || !q->jx) return 0;
    for (int64_t k = 0; k < q->nh; k++) {
        int64_t h = q->holes[k];
        if (q->jx[h] != 0 || q->jy[h] != 0 || q->jz[h] != 0) return 1;
    }
    return 0;
}

/* x^1.5 for dragBack's weights.  The reference calls math.Pow(x, 1.5), which
 * Go evaluates as Exp(0.5 Log x) times the binary fraction of x (a few ulp
 * from x^1.5); restated here as x * sqrt(x).  Only the ratios w_h / sum(w)
 * are used, so with one spinning hole the share is w / w = 1 either way. */
static double pow15(double x) { return x * sqrt(x); }

/* gravitomagnetic, framedrag.go:34-54: B at body i from every spinning alive
 * hole except i, softened like the Newtonian pull */
static void gravitomagnetic(const split_ctx *q, int64_t i, double b[3]) {
    double bx = 0, by = 0, bz = 0;
    for (int64_t k = 0; k < q->nh; k++) {
        int64_t h = q->holes[k];
        if (h == i || !q->alive[h]) continue;
        double jx = q->jx[h], jy = q->jy[h], jz = q->jz[h];
        if (jx * jx + jy * jy + jz * jz == 0) continue;
        double rx = q->px[i] - q->px[h], ry = q->py[i] - q->py[h], rz = q->pz[i] - q->pz[h];
        double r2 = (rx * rx + ry * ry + rz * rz) + q->eps2;
        double inv = 1 / sqrt(r2);
        double k3 = 2 / (q->c * q->c) * inv * inv * inv;
        double sc = 3 * (jx * rx + jy * ry + jz * rz) / r2;
        bx = bx + (jx - rx * sc) * k3;
        by = by + (jy - ry * sc) * k3;
        bz = bz + (jz - rz * sc) * k3;
    }
    b[0] = bx; b[1] = by; b[2] = bz;
}

/* dragBack, framedrag.go:74-113: the holes take -m du in proportion to their
 * weights |J| / r^3, as momentum and as spin (-(r x m du f)) */
static void drag_back(const split_ctx *q, int64_t i, double dux, double duy, double duz) {
    double mi = q->mass[i];
    double pxm = dux * mi, pym = duy * mi, pzm = duz * mi;
    double total = 0;
    for (int64_t k = 0; k < q->nh; k++) {
        int64_t h = q->holes[k];
        double w = 0;
        if (h != i && q->alive[h]) {
            double rx = q->px[i] - q->px[h], ry = q->py[i] - q->py[h], rz = q->pz[i] - q->pz[h];
            double j = sqrt(q->jx[h] * q->jx[h] + q->jy[h] * q->jy[h] + q->jz[h] * q->jz[h]);
            if (j > 0) w = j / pow15(rx * rx + ry * ry + rz * rz);
        }
        q->weights[k] = w;
        total += w;
    }
    if (total == 0) return;
    for (int64_t k = 0; k < q->nh; k++) {
        if (q->weights[k] == 0) continue;
        int64_t h = q->holes[k];
        double f = q->weights[k] / total, mh = q->mass[h];
        double phx = pxm * -f, phy = pym * -f, phz = pzm * -f;
        q->vx[h] += phx / mh;
        q->vy[h] += phy / mh;
        q->vz[h] += phz / mh;
        double rx = q->px[i] - q->px[h], ry = q->py[i] - q->py[h], rz = q->pz[i] - q->pz[h];
        double gx = pxm * f, gy = pym * f, gz = pzm * f;
        q->jx[h] += (ry * gz - rz * gy) * -1;
        q->jy[h] += (rz * gx - rx * gz) * -1;
        q->jz[h] += (rx * gy - ry * gx) * -1;
    }
}

/* dragKick, framedrag.go:56-72: Boris rotation of u by B over time dt, then
 * dragBack */
static void drag_kick(const split_ctx *q, int64_t i, double dt) {
    double b[3];
    gravitomagnetic(q, i, b);
    if (b[0] * b[0] + b[1] * b[1] + b[2] * b[2] == 0) return;
    double ux = q->vx[i], uy = q->vy[i], uz = q->vz[i];
    double g = gamma_of(ux, uy, uz, q->c);
    double hs = 0.5 * dt / g;
    double tx = b[0] * hs, ty = b[1] * hs, tz = b[2] * hs;
    double upx = ux + (uy * tz - uz * ty), upy = uy + (uz * tx - ux * tz), upz = uz + (ux * ty - uy * tx);
    double s2 = 2 / (1 + (tx * tx + ty * ty + tz * tz));
    double sx = tx * s2, sy = ty * s2, sz = tz * s2;
    double unx = ux + (upy * sz - upz * sy), uny = uy + (upz * sx - upx * sz), unz = uz + (upx * sy - upy * sx);
    q->vx[i] = unx;
    q->vy[i] = uny;
    q->vz[i] = unz;
    drag_back(q, i, unx - ux, uny - uy, unz - uz);
}

/* dragOuter, framedrag.go:115-127: every alive body that is neither a hole
 * nor near, in state order */
static void drag_outer(const split_ctx *q, double dt) {
    if (!spinning(q)) return;
    for (int64_t i = 0; i < q->n; i++)
        if (q->kind[i] != KIND_HOLE && !q->near[i] && q->alive[i]) drag_kick(q, i, dt);
}

/* dragFast, framedrag.go:129-143: the near list, then the holes */
static void drag_fast(const split_ctx *q, double dt) {
    if (!spinning(q)) return;
    for (int64_t j = 0; j < q->nn; j++) drag_kick(q, q->nearList[j], dt);
    for (int64_t k = 0; k < q->nh; k++) drag_kick(q, q->holes[k], dt);
}

/* Test hook: one dragKick of body i against the alive holes of the state
 * (framedrag_test.go:82-89 calls dragKick directly). */
void or_drag_kick(double *px, double *py, double *pz, double *vx, double *vy, double *vz, const double *mass,
                  const uint8_t *kind, const uint8_t *alive, double *jx, double *jy, double *jz, int64_t n, int64_t i,
                  double dt, double eps2, double c) {
    split_ctx q = {px, py, pz, vx, vy, vz, mass, kind, n, eps2, c, 0, NULL, 0, NULL, NULL, 0, alive, jx, jy, jz,
                   NULL, NULL, GRAVITY_INSTANT, 0, 0};
    int32_t *holes = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    q.weights = malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int64_t k = 0; k < n; k++)
        if (alive[k] && kind[k] == KIND_HOLE) holes[q.nh++] = (int32_t)k;
    q.holes = holes;
    drag_kick(&q, i, dt);
    free(holes);
    free(q.weights);
}

/* Test hooks: dragOuter over time dt with the near flags near (NULL: none),
 * and dragFast over the near list list[0 .. count). */
void or_drag_pass(double *px, double *py, double *pz, double *vx, double *vy, double *vz, const double *mass,
                  const uint8_t *kind, const uint8_t *alive, double *jx, double *jy, double *jz, int64_t n,
                  const uint8_t *near, const int32_t *list, int64_t count, int fast, double dt, double eps2,
                  double c) {
    split_ctx q = {px, py, pz, vx, vy, vz, mass, kind, n, eps2, c, 0, NULL, 0, NULL, NULL, 0, alive, jx, jy, jz,
                   NULL, NULL, GRAVITY_INSTANT, 0, 0};
    int32_t *holes = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    q.weights = malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    q.near = calloc((size_t)(n > 0 ? n : 1), 1);
    if (near) memcpy(q.near, near, (size_t)n);
    q.nearList = (int32_t *)list;
    q.nn = count;
    for (int64_t k = 0; k < n; k++)
        if (alive[k] && kind[k] == KIND_HOLE) holes[q.nh++] = (int32_t)k;
    q.holes = holes;
    if (fast)
        drag_fast(&q, dt);
    else
        drag_outer(&q, dt);
    free(holes);
    free(q.weights);
    free(q.near);
}

/* stepSplit, split.go:164-199, for a state with alive holes and no gas; the
 * slow forces are the tree's (theta, eps, dm_off).  jx, jy, jz (the hole
 * spins, may be NULL) switch on frame dragging (framedrag.go) when c > 0:
 * after the first outer kick and before the last one, and around the drifts
 * of every substep.  Returns the substeps. */
int or_split_step(or_solver *s, double *px, double *py, double *pz, double *vx, double *vy, double *vz,
                  const double *mass, const uint8_t *kind, const uint8_t *alive, double *jx, double *jy, double *jz,
                  int64_t n, double theta, double eps, int dm_off, double dt, double c, int acc_valid, double *ax,
                  double *ay, double *az, or_retard *r, const uint64_t *ids, double time, int mode) {
    split_ctx q = {px, py, pz, vx, vy, vz, mass, kind, n, eps * eps, c, dm_off, NULL, 0, NULL, NULL, 0, alive,
                   jx, jy, jz, NULL, NULL, r ? mode : GRAVITY_INSTANT, time, dt};
    int32_t *holes = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    q.weights = malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    q.near = malloc((size_t)(n > 0 ? n : 1));
    q.nearList = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    double *farx = malloc((size_t)(n > 0 ? n : 1) * 8 * 6);
    double *fary = farx + n, *farz = fary + n, *fsx = farz + n, *fsy = fsx + n, *fsz = fsy + n;
    for (int64_t i = 0; i < n; i++)
        if (alive[i] && kind[i] == KIND_HOLE) holes[q.nh++] = (int32_t)i; /* collectHoles, sim.go:291-299 */
    q.holes = holes;
    q.hists = calloc((size_t)(q.nh > 0 ? q.nh : 1), sizeof(or_hist *));
    if (r && mode != GRAVITY_INSTANT && c > 0) /* bindHistories, retarded.go:106-116 */
        for (int64_t k = 0; k < q.nh; k++) q.hists[k] = hist_find(r, body_id(ids, holes[k]));
    theta = theta == 0 ? DEFAULT_THETA : theta;
    if (!acc_valid) or_accelerations(s, px, py, pz, mass, kind, alive, n, theta, eps, dm_off, s->monopole, s->workers,
                                     ax, ay, az);
    if (q.nh == 0) {
        /* step() without holes (sim.go:266-277): kick, drift (:316-343), forces, kick */
        or_kick(n, vx, vy, vz, ax, ay, az, 0.5 * dt);
        if (c <= 0) {
            or_drift(n, px, py, pz, vx, vy, vz, dt);
        } else {
            double invC2 = 1 / (c * c);
            for (int64_t i = 0; i < n; i++) {
                double ux = vx[i], uy = vy[i], uz = vz[i];
                double f = dt / sqrt(1 + (ux * ux + uy * uy + uz * uz) * invC2);
                px[i] += f * ux;
                py[i] += f * uy;
                pz[i] += f * uz;
            }
        }
        or_accelerations(s, px, py, pz, mass, kind, alive, n, theta, eps, dm_off, s->monopole, s->workers, ax, ay, az);
        or_kick(n, vx, vy, vz, ax, ay, az, 0.5 * dt);
        if (r) record_holes(r, px, py, pz, vx, vy, vz, kind, alive, ids, n, time + dt, dt, c, mode);
        free(holes);
        free(q.hists);
        free(q.weights);
        free(q.near);
        free(q.nearList);
        free(farx);
        return 1;
    }
    int sub = classify(&q, dt);
    outer_hole_forces(&q, farx, fary, farz);
    for (int64_t i = 0; i < n; i++) { /* kickOuter, split.go:201-208 */
        vx[i] += 0.5 * dt * (ax[i] + farx[i]);
        vy[i] += 0.5 * dt * (ay[i] + fary[i]);
        vz[i] += 0.5 * dt * (az[i] + farz[i]);
    }
    drag_outer(&q, 0.5 * dt);
    double h = dt / (double)sub;
    fast_forces(&q, fsx, fsy, fsz);
    for (int k = 0; k < sub; k++) {
        kick_idx(&q, q.nearList, q.nn, fsx, fsy, fsz, 0.5 * h);
        kick_idx(&q, holes, q.nh, fsx, fsy, fsz, 0.5 * h);
        drag_fast(&q, 0.5 * h);
        drift_idx(&q, q.nearList, q.nn, h);
        drift_idx(&q, holes, q.nh, h);
        fast_forces(&q, fsx, fsy, fsz);
        drag_fast(&q, 0.5 * h);
        kick_idx(&q, q.nearList, q.nn, fsx, fsy, fsz, 0.5 * h);
        kick_idx(&q, holes, q.nh, fsx, fsy, fsz, 0.5 * h);
    }
    for (int64_t i = 0; i < n; i++) { /* driftFar, split.go:225-238 */
        if (kind[i] == KIND_HOLE || q.near[i]) continue;
        double f = dt / gamma_of(vx[i], vy[i], vz[i], c);
        px[i] += f * vx[i];
        py[i] += f * vy[i];
        pz[i] += f * vz[i];
    }
    or_accelerations(s, px, py, pz, mass, kind, alive, n, theta, eps, dm_off, s->monopole, s->workers, ax, ay, az);
    outer_hole_forces(&q, farx, fary, farz);
    drag_outer(&q, 0.5 * dt);
    for (int64_t i = 0; i < n; i++) {
        vx[i] += 0.5 * dt * (ax[i] + farx[i]);
        vy[i] += 0.5 * dt * (ay[i] + fary[i]);
        vz[i] += 0.5 * dt * (az[i] + farz[i]);
    }
    if (r) record_holes(r, px, py, pz, vx, vy, vz, kind, alive, ids, n, time + dt, dt, c, mode); /* sim.go:285-287 */
    free(holes);
    free(q.hists);
    free(q.weights);
    free(q.near);
    free(q.nearList);
    free(farx);
    return sub;
}

/* Test hook: sourcePos (retarded.go:118-155) of alive hole number k (in
 * state order) for an observer at (x, y, z), with the histories of r bound
 * as in a step at the given time. */
int or_source_pos(or_retard *r, double *px, double *py, double *pz, const uint8_t *kind, const uint8_t *alive,
                  const uint64_t *ids, int64_t n, int64_t k, double x, double y, double z, double time, double dt,
                  double c, int mode, double out[3]) {
    split_ctx q = {px, py, pz, NULL, NULL, NULL, NULL, kind, n, 0, c, 0, NULL, 0, NULL, NULL, 0, alive, NULL, NULL,
                   NULL, NULL, NULL, mode, time, dt};
    int32_t *holes = malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    for (int64_t i = 0; i < n; i++)
        if (alive[i] && kind[i] == KIND_HOLE) holes[q.nh++] = (int32_t)i;
    q.holes = holes;
    q.hists = calloc((size_t)(q.nh > 0 ? q.nh : 1), sizeof(or_hist *));
    for (int64_t j = 0; j < q.nh; j++) q.hists[j] = hist_find(r, body_id(ids, holes[j]));
    int vis = k < q.nh ? source_pos(&q, k, x, y, z, out) : 0;
    free(holes);
    free(q.hists);
    return vis;
}

/* ------------------------------------------------------------------ */
/* Accessors for the Python wrapper                                     */
/* ------------------------------------------------------------------ */
int64_t or_n_stars(const or_solver *s) { return s->n_stars; }
int64_t or_n_nodes(const or_solver *s) { return s->n_nodes; }
const or_node *or_nodes(const or_solver *s) { return s->nodes; }
const uint64_t *or_keys(const or_solver *s) { return s->keys; }
const int32_t *or_order(const or_solver *s) { return s->order; }
const double *or_sorted(const or_solver *s, int which) {
    return which == 0 ? s->sx : which == 1 ? s->sy : which == 2 ? s->sz : s->sm;
}
void or_set_monopole(or_solver *s, int monopole) { s->monopole = monopole; }
void or_set_workers(or_solver *s, int workers) { s->workers = workers; }
int or_node_size(void) { return (int)sizeof(or_node); }

/* The pair model (pair.go:13-47, relativity.go:49-54) for the tests that
 * restate internal/sim/pair_test.go. */
double or_pair_force(double m, double r2, double eps2, double rs) { return pair_force(m, r2, eps2, rs); }
double or_pair_potential(double mi, double mj, double r2, double eps2, double rs) {
    return pair_potential(mi, mj, r2, eps2, rs);
}
double or_pair_rs(int ki, double mi, int kj, double mj, double c) { return pair_rs(ki, mi, kj, mj, c); }
