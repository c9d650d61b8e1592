/*
 * sph_oracle.c -- CPU restatement of the gravitas SPH neighbour tree and the
 * hydro force evaluation that uses it.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP hydro backend is checked
 * against (tests/ and bench.py's cpu_baseline leg).  No product path links,
 * loads or calls it.
 *
 * Restated from the Go reference (paths relative to the gravitas repository),
 * in the same floating-point order:
 *   tree.build / keys / radixSort   internal/sph/tree.go:49-124 (any stable
 *                                   sort by key gives the same permutation)
 *   buildNode / split / refit       internal/sph/tree.go:126-180
 *   refitH                          internal/sph/tree.go:184-201
 *   each                            internal/sph/tree.go:205-239
 *   candidates                      internal/sph/tree.go:244-267
 *   scan                            internal/sph/tree.go:273-309
 *   sortLeaves                      internal/sph/tree.go:314-331
 *   boxBoxDist2 / boxDist2          internal/sph/tree.go:334-355
 *   W / DW                          internal/sph/kernel.go:9-36
 *   Solver.Forces, initH            internal/sph/sph.go:83-263, :277-302
 *   lowerWake                       internal/sph/sph.go:266-274
 *
 * math.Cbrt is Go's own routine, whose source is not available here.
 * or_cbrt is this project's cube root: range reduction to [0.5, 4), Newton
 * steps, and a last Newton step on the exactly computed residual y^3 - m (FMA
 * splits).  It is correctly rounded on every input the tests check (exact
 * rational comparison), so it agrees with any correctly rounded cube root.
 * Go documents its Cbrt as accurate to within 1 ulp; wherever Go's result is
 * the correctly rounded one, the two are equal.  The HIP backend uses the
 * same sequence of IEEE operations, so the oracle and the GPU agree bit for
 * bit.
 *
 * Build with -O2 -ffp-contract=off (oracle/Makefile), like the gravity oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Go's math.Min / math.Max and the builtin min / max on floats: NaN
 * propagates, -0 < +0. */
static double gmin(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && b == 0) return signbit(a) ? a : b;
    return a < b ? a : b;
}
static double gmax(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && b == 0) return signbit(a) ? b : a;
    return a > b ? a : b;
}

static const double kPi = 3.14159265358979323846; /* math.Pi as float64 */

/* ------------------------------------------------------------------ */
/* Cube root                                                           */

double or_cbrt(double x) {
    if (!(x != 0) || isinf(x) || isnan(x)) return x;
    int neg = x < 0;
    double a = neg ? -x : x;
    int e;
    double f = frexp(a, &e); /* a = f 2^e, f in [0.5, 1) */
    int q = e >= 0 ? e / 3 : -((-e + 2) / 3);
    int r = e - 3 * q;           /* 0, 1 or 2 */
    double m = ldexp(f, r);      /* m in [0.5, 4), exact */
    /* linear start, then Newton y <- y - (y^3 - m) / (3 y^2) */
    double y = 0.6 + 0.25 * m;
    for (int it = 0; it < 5; it++) y = y - (y * y * y - m) / (3 * (y * y));
    /* y^3 - m with the products split exactly by fma */
    double t = y * y, tl = fma(y, y, -t);
    double p = t * y, pl = fma(t, y, -p);
    double res = (p - m) + (pl + tl * y);
    y = y - res / (3 * (y * y));
    y = ldexp(y, q);
    return neg ? -y : y;
}

/* ------------------------------------------------------------------ */
/* Kernel: kernel.go:9-36                                              */

double or_sph_w(double r, double h) {
    double q = r / h;
    double sigma = 1 / (kPi * h * h * h);
    if (q < 1) return sigma * (1 - 1.5 * q * q + 0.75 * q * q * q);
    if (q < 2) {
        double t = 2 - q;
        return sigma * 0.25 * t * t * t;
    }
    return 0;
}

double or_sph_dw(double r, double h) {
    double q = r / h;
    double sigma = 1 / (kPi * h * h * h * h);
    if (q < 1) return sigma * (-3 * q + 2.25 * q * q);
    if (q < 2) {
        double t = 2 - q;
        return -sigma * 0.75 * t * t;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Tree: tree.go                                                       */

typedef struct {
    double minx, miny, minz, maxx, maxy, maxz, hmax;
    int32_t lo, hi, left, right;
} sph_node;

typedef struct {
    int32_t n;
    uint64_t *keys;
    int32_t *order, *pos;
    double *sx, *sy, *sz, *sh;
    sph_node *nodes;
    int32_t nnodes, cap;
    int32_t *leaves;
    int32_t nleaves;
    int32_t *stack;
    int32_t stackCap;
} sph_tree;

void *or_sph_tree_new(void) { return calloc(1, sizeof(sph_tree)); }

void or_sph_tree_free(void *p) {
    sph_tree *t = (sph_tree *)p;
    if (!t) return;
    free(t->keys); free(t->order); free(t->pos);
    free(t->sx); free(t->sy); free(t->sz); free(t->sh);
    free(t->nodes); free(t->leaves); free(t->stack);
    free(t);
}

/* The 21 low bits of v spread to every third bit (bit b -> bit 3b). */
static uint64_t spread21(uint64_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 21; b++) out |= ((v >> b) & 1) << (3 * b);
    return out;
}

static const uint64_t *g_keys; /* qsort context */
static const int32_t *g_pos;
static int by_key_then_pos(const void *a, const void *b) {
    int32_t i = *(const int32_t *)a, j = *(const int32_t *)b;
    if (g_keys[i] != g_keys[j]) return g_keys[i] < g_keys[j] ? -1 : 1;
    return g_pos[i] < g_pos[j] ? -1 : (g_pos[i] > g_pos[j]);
}

static int32_t push_node(sph_tree *t, int32_t lo, int32_t hi) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 64;
        t->nodes = (sph_node *)realloc(t->nodes, sizeof(sph_node) * t->cap);
    }
    sph_node *nd = &t->nodes[t->nnodes];
    memset(nd, 0, sizeof *nd);
    nd->lo = lo; nd->hi = hi; nd->left = -1; nd->right = -1;
    return t->nnodes++;
}

/* refit, tree.go:164-180 */
static void refit(sph_tree *t, int32_t ni) {
    sph_node *nd = &t->nodes[ni];
    nd->minx = nd->miny = nd->minz = INFINITY;
    nd->maxx = nd->maxy = nd->maxz = -INFINITY;
    nd->hmax = 0;
    if (nd->left < 0) {
        for (int32_t k = nd->lo; k < nd->hi; k++) {
            nd->minx = gmin(nd->minx, t->sx[k]); nd->maxx = gmax(nd->maxx, t->sx[k]);
            nd->miny = gmin(nd->miny, t->sy[k]); nd->maxy = gmax(nd->maxy, t->sy[k]);
            nd->minz = gmin(nd->minz, t->sz[k]); nd->maxz = gmax(nd->maxz, t->sz[k]);
            nd->hmax = gmax(nd->hmax, t->sh[k]);
        }
        return;
    }
    const sph_node *c[2] = {&t->nodes[nd->left], &t->nodes[nd->right]};
    for (int s = 0; s < 2; s++) {
        nd->minx = gmin(nd->minx, c[s]->minx); nd->maxx = gmax(nd->maxx, c[s]->maxx);
        nd->miny = gmin(nd->miny, c[s]->miny); nd->maxy = gmax(nd->maxy, c[s]->maxy);
        nd->minz = gmin(nd->minz, c[s]->minz); nd->maxz = gmax(nd->maxz, c[s]->maxz);
        nd->hmax = gmax(nd->hmax, c[s]->hmax);
    }
}

/* split, tree.go:150-161: at the highest key bit in which the range differs,
 * or in the middle of a run of equal keys. */
static int32_t split(const sph_tree *t, int32_t lo, int32_t hi) {
    uint64_t a = t->keys[lo], b = t->keys[hi - 1];
    if (a == b) return (lo + hi) / 2;
    uint64_t bit = (uint64_t)1 << (63 - __builtin_clzll(a ^ b));
    int32_t l = 0, h = hi - lo; /* first k with the bit set (sort.Search) */
    while (l < h) {
        int32_t mid = l + (h - l) / 2;
        if (t->keys[lo + mid] & bit) h = mid; else l = mid + 1;
    }
    return lo + l;
}

#define SPH_LEAF 8

/* buildNode, tree.go:126-146 (pre-order, left first) */
static int32_t build_node(sph_tree *t, int32_t lo, int32_t hi) {
    int32_t ni = push_node(t, lo, hi);
    if (hi - lo <= SPH_LEAF) {
        refit(t, ni);
        t->leaves[t->nleaves++] = ni;
        return ni;
    }
    int32_t mid = split(t, lo, hi);
    int32_t l = build_node(t, lo, mid);
    int32_t r = build_node(t, mid, hi);
    t->nodes[ni].left = l;
    t->nodes[ni].right = r;
    refit(t, ni);
    return ni;
}

/* tree.build, tree.go:91-124.  gas holds body indices; x, y, z, h are
 * indexed by body. */
int32_t or_sph_tree_build(void *p, const int32_t *gas, int32_t n, const double *x, const double *y,
                          const double *z, const double *h) {
    sph_tree *t = (sph_tree *)p;
    double minX = INFINITY, minY = INFINITY, minZ = INFINITY;
    double maxX = -INFINITY, maxY = -INFINITY, maxZ = -INFINITY;
    for (int32_t k = 0; k < n; k++) {
        int32_t i = gas[k];
        minX = gmin(minX, x[i]); maxX = gmax(maxX, x[i]);
        minY = gmin(minY, y[i]); maxY = gmax(maxY, y[i]);
        minZ = gmin(minZ, z[i]); maxZ = gmax(maxZ, z[i]);
    }
    double side = gmax(gmax(maxX - minX, maxY - minY), gmax(maxZ - minZ, 1e-12)) * (1 + 1e-9);
    double scale = (double)((1 << 21) - 1) / side;
    t->n = n;
    t->keys = (uint64_t *)realloc(t->keys, sizeof(uint64_t) * (n ? n : 1));
    t->order = (int32_t *)realloc(t->order, sizeof(int32_t) * (n ? n : 1));
    t->pos = (int32_t *)realloc(t->pos, sizeof(int32_t) * (n ? n : 1));
    uint64_t *bykey = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    for (int32_t k = 0; k < n; k++) {
        int32_t i = gas[k];
        uint64_t kx = spread21((uint64_t)((x[i] - minX) * scale));
        uint64_t ky = spread21((uint64_t)((y[i] - minY) * scale));
        uint64_t kz = spread21((uint64_t)((z[i] - minZ) * scale));
        bykey[k] = kx << 2 | ky << 1 | kz;
        t->pos[k] = k;
        idx[k] = k;
    }
    g_keys = bykey; g_pos = t->pos;
    qsort(idx, n, sizeof(int32_t), by_key_then_pos); /* stable by construction: ties by gas position */
    for (int32_t k = 0; k < n; k++) {
        t->keys[k] = bykey[idx[k]];
        t->order[k] = gas[idx[k]];
    }
    free(bykey); free(idx);
    size_t sz = sizeof(double) * (n ? n : 1);
    t->sx = (double *)realloc(t->sx, sz); t->sy = (double *)realloc(t->sy, sz);
    t->sz = (double *)realloc(t->sz, sz); t->sh = (double *)realloc(t->sh, sz);
    for (int32_t k = 0; k < n; k++) {
        int32_t i = t->order[k];
        t->sx[k] = x[i]; t->sy[k] = y[i]; t->sz[k] = z[i]; t->sh[k] = h[i];
    }
    t->nnodes = 0;
    t->nleaves = 0;
    t->leaves = (int32_t *)realloc(t->leaves, sizeof(int32_t) * (n ? n : 1));
    if (n > 0) build_node(t, 0, n);
    return t->nnodes;
}

/* refitH, tree.go:184-201 */
void or_sph_tree_refit_h(void *p, const double *h) {
    sph_tree *t = (sph_tree *)p;
    for (int32_t k = 0; k < t->n; k++) t->sh[k] = h[t->order[k]];
    for (int32_t ni = t->nnodes - 1; ni >= 0; ni--) {
        sph_node *nd = &t->nodes[ni];
        if (nd->left < 0) {
            nd->hmax = 0;
            for (int32_t k = nd->lo; k < nd->hi; k++) nd->hmax = gmax(nd->hmax, h[t->order[k]]);
        } else {
            nd->hmax = gmax(t->nodes[nd->left].hmax, t->nodes[nd->right].hmax);
        }
    }
}

/* sortLeaves, tree.go:314-331: every leaf by decreasing h (stable). */
void or_sph_tree_sort_leaves(void *p) {
    sph_tree *t = (sph_tree *)p;
    for (int32_t l = 0; l < t->nleaves; l++) {
        const sph_node *nd = &t->nodes[t->leaves[l]];
        for (int32_t a = nd->lo + 1; a < nd->hi; a++) {
            for (int32_t b = a; b > nd->lo && t->sh[b] > t->sh[b - 1]; b--) {
                int32_t ti = t->order[b]; t->order[b] = t->order[b - 1]; t->order[b - 1] = ti;
                double v;
                v = t->sx[b]; t->sx[b] = t->sx[b - 1]; t->sx[b - 1] = v;
                v = t->sy[b]; t->sy[b] = t->sy[b - 1]; t->sy[b - 1] = v;
                v = t->sz[b]; t->sz[b] = t->sz[b - 1]; t->sz[b - 1] = v;
                v = t->sh[b]; t->sh[b] = t->sh[b - 1]; t->sh[b - 1] = v;
            }
        }
    }
}

int32_t or_sph_tree_nnodes(const void *p) { return ((const sph_tree *)p)->nnodes; }
const sph_node *or_sph_tree_nodes(const void *p) { return ((const sph_tree *)p)->nodes; }
const uint64_t *or_sph_tree_keys(const void *p) { return ((const sph_tree *)p)->keys; }
const int32_t *or_sph_tree_order(const void *p) { return ((const sph_tree *)p)->order; }
int32_t or_sph_tree_nleaves(const void *p) { return ((const sph_tree *)p)->nleaves; }
const int32_t *or_sph_tree_leaves(const void *p) { return ((const sph_tree *)p)->leaves; }

/* boxDist2, tree.go:349-355 */
static double box_dist2(const sph_node *nd, double px, double py, double pz) {
    double dx = gmax(gmax(nd->minx - px, px - nd->maxx), 0);
    double dy = gmax(gmax(nd->miny - py, py - nd->maxy), 0);
    double dz = gmax(gmax(nd->minz - pz, pz - nd->maxz), 0);
    return dx * dx + dy * dy + dz * dz;
}

/* boxBoxDist2, tree.go:334-346 */
static double box_box_dist2(const sph_node *a, const sph_node *b) {
    double d[3] = {gmax(gmax(a->minx - b->maxx, b->minx - a->maxx), 0),
                   gmax(gmax(a->miny - b->maxy, b->miny - a->maxy), 0),
                   gmax(gmax(a->minz - b->maxz, b->minz - a->maxz), 0)};
    double d2 = 0;
    for (int c = 0; c < 3; c++) d2 += d[c] * d[c];
    return d2;
}

static void stack_push(sph_tree *t, int32_t *top, int32_t v) {
    if (*top == t->stackCap) {
        t->stackCap = t->stackCap ? 2 * t->stackCap : 64;
        t->stack = (int32_t *)realloc(t->stack, sizeof(int32_t) * t->stackCap);
    }
    t->stack[(*top)++] = v;
}

typedef void (*visit_fn)(void *ctx, int32_t j, double r2);

/* each, tree.go:205-239 */
static void each(sph_tree *t, double px, double py, double pz, double radius, int sym, visit_fn fn, void *ctx) {
    if (t->nnodes == 0) return;
    int32_t top = 0;
    stack_push(t, &top, 0);
    while (top > 0) {
        const sph_node *nd = &t->nodes[t->stack[--top]];
        double reach = radius;
        if (sym && 2 * nd->hmax > reach) reach = 2 * nd->hmax;
        if (box_dist2(nd, px, py, pz) >= reach * reach) continue;
        if (nd->left >= 0) {
            int32_t l = nd->left, r = nd->right;
            stack_push(t, &top, l);
            stack_push(t, &top, r);
            continue;
        }
        for (int32_t k = nd->lo; k < nd->hi; k++) {
            double ex = t->sx[k] - px, ey = t->sy[k] - py, ez = t->sz[k] - pz;
            double r2 = ex * ex + ey * ey + ez * ez;
            double lim = radius;
            if (sym && 2 * t->sh[k] > lim) lim = 2 * t->sh[k];
            if (r2 < lim * lim) fn(ctx, t->order[k], r2);
        }
    }
}

/* candidates, tree.go:244-267: the leaves a group walk from leaf must scan. */
static int32_t candidates(sph_tree *t, int32_t leaf, int sym, int32_t **cand, int32_t *candCap) {
    const sph_node *lf = &t->nodes[leaf];
    int32_t nc = 0, top = 0;
    stack_push(t, &top, 0);
    while (top > 0) {
        int32_t ni = t->stack[--top];
        const sph_node *nd = &t->nodes[ni];
        double reach = 2 * lf->hmax;
        if (sym && 2 * nd->hmax > reach) reach = 2 * nd->hmax;
        if (box_box_dist2(nd, lf) >= reach * reach) continue;
        if (nd->left >= 0) {
            int32_t l = nd->left, r = nd->right;
            stack_push(t, &top, l);
            stack_push(t, &top, r);
            continue;
        }
        if (nc == *candCap) {
            *candCap = *candCap ? 2 * *candCap : 64;
            *cand = (int32_t *)realloc(*cand, sizeof(int32_t) * *candCap);
        }
        (*cand)[nc++] = ni;
    }
    return nc;
}

/* scan, tree.go:273-309 */
static void scan(const sph_tree *t, const int32_t *cand, int32_t nc, double px, double py, double pz,
                 double radius, int sym, visit_fn fn, void *ctx) {
    double r2max = radius * radius;
    for (int32_t c = 0; c < nc; c++) {
        const sph_node *nd = &t->nodes[cand[c]];
        double reach2 = r2max;
        if (sym) {
            double l = 2 * nd->hmax;
            if (l * l > reach2) reach2 = l * l;
        }
        double d2 = box_dist2(nd, px, py, pz);
        if (d2 >= reach2) continue;
        int scatterOnly = sym && d2 >= r2max;
        for (int32_t k = nd->lo; k < nd->hi; k++) {
            if (scatterOnly && 4 * t->sh[k] * t->sh[k] <= d2) break;
            double ex = t->sx[k] - px, ey = t->sy[k] - py, ez = t->sz[k] - pz;
            double r2 = ex * ex + ey * ey + ez * ez;
            double lim2 = r2max;
            if (sym) {
                double l = 2 * t->sh[k];
                if (l * l > lim2) lim2 = l * l;
            }
            if (r2 < lim2) fn(ctx, t->order[k], r2);
        }
    }
}

/* Test hooks: the bodies each() and a leaf's group scan report, in order. */
typedef struct {
    int32_t *out;
    double *r2;
    int32_t n, max;
} collect_ctx;

static void collect(void *c, int32_t j, double r2) {
    collect_ctx *cc = (collect_ctx *)c;
    if (cc->n < cc->max) {
        cc->out[cc->n] = j;
        if (cc->r2) cc->r2[cc->n] = r2;
    }
    cc->n++;
}

int32_t or_sph_each(void *p, double px, double py, double pz, double radius, int sym, int32_t *out, double *r2,
                    int32_t max) {
    collect_ctx cc = {out, r2, 0, max};
    each((sph_tree *)p, px, py, pz, radius, sym, collect, &cc);
    return cc.n;
}

int32_t or_sph_group_scan(void *p, int32_t leaf, int sym, double px, double py, double pz, double radius,
                          int32_t *out, double *r2, int32_t max) {
    sph_tree *t = (sph_tree *)p;
    int32_t *cand = NULL, cap = 0;
    int32_t nc = candidates(t, leaf, sym, &cand, &cap);
    collect_ctx cc = {out, r2, 0, max};
    scan(t, cand, nc, px, py, pz, radius, sym, collect, &cc);
    free(cand);
    return cc.n;
}

/* ------------------------------------------------------------------ */
/* Solver.Forces, sph.go:83-263                                        */

#define MAX_DENSITY_ITERATIONS 6
#define H_TOLERANCE 0.05
#define BETA_FACTOR 2
#define H_MIN_SOFTENING 0.25
#define COURANT 0.3
#define LIMITER_RATIO 4

typedef struct {
    const double *m;
    double hi, rho;
} density_ctx;

static void density_term(void *c, int32_t j, double r2) {
    density_ctx *d = (density_ctx *)c;
    d->rho += d->m[j] * or_sph_w(sqrt(r2), d->hi);
}

typedef struct {
    const double *x, *y, *z, *vx, *vy, *vz, *m, *h, *rho, *P, *C, *step;
    int32_t i;
    double hi, rhoi, fi, alpha;
    double ax, ay, az, du, dv, vsig, minStep, maxStep;
} force_ctx;

static void force_term(void *c, int32_t j, double r2) {
    force_ctx *f = (force_ctx *)c;
    int32_t i = f->i;
    if (j == i || r2 == 0) return;
    if (f->step && f->step[j] > 0) {
        f->minStep = gmin(f->minStep, f->step[j]);
        f->maxStep = gmax(f->maxStep, f->step[j]);
    }
    double r = sqrt(r2);
    double ex = (f->x[i] - f->x[j]) / r, ey = (f->y[i] - f->y[j]) / r, ez = (f->z[i] - f->z[j]) / r;
    double vx = f->vx[i] - f->vx[j], vy = f->vy[i] - f->vy[j], vz = f->vz[i] - f->vz[j];
    double vr = vx * ex + vy * ey + vz * ez;
    f->vsig = gmax(f->vsig, f->C[i] + f->C[j] - 3 * gmin(0, vr));
    double dwi = or_sph_dw(r, f->hi), dwj = or_sph_dw(r, f->h[j]);
    double dwm = 0.5 * (dwi + dwj);
    double fj = f->P[j] / (f->rho[j] * f->rho[j]);
    double visc = 0;
    if (vr < 0) {
        double hm = 0.5 * (f->hi + f->h[j]);
        double mu = hm * vr * r / (r2 + 0.01 * hm * hm);
        visc = (-f->alpha * 0.5 * (f->C[i] + f->C[j]) * mu + BETA_FACTOR * f->alpha * mu * mu) /
               (0.5 * (f->rhoi + f->rho[j]));
    }
    double fo = -f->m[j] * (f->fi * dwi + fj * dwj + visc * dwm);
    f->ax = f->ax + fo * ex;
    f->ay = f->ay + fo * ey;
    f->az = f->az + fo * ez;
    f->du += f->m[j] * (f->fi * dwi + 0.5 * visc * dwm) * vr;
    f->dv -= f->m[j] * dwi * vr;
}

/* lowerWake, sph.go:266-274 (sequential here: min of the float values) */
typedef struct {
    const double *step;
    uint64_t *wake;
    double dt;
} wake_ctx;

static void wake_term(void *c, int32_t j, double r2) {
    (void)r2;
    wake_ctx *w = (wake_ctx *)c;
    if (w->step[j] > LIMITER_RATIO * w->dt) {
        double old;
        memcpy(&old, &w->wake[j], sizeof old);
        if (!(old <= w->dt)) memcpy(&w->wake[j], &w->dt, sizeof w->dt);
    }
}

/* setH, sph.go:121-131: store the density, move h towards its target and
 * report whether it must iterate again. */
static int set_h(int32_t i, double r, int last, double hMin, double ngb, const double *m, double *h,
                 double *rho) {
    double hi = h[i];
    rho[i] = r;
    double target = gmax(hMin, 0.5 * or_cbrt(3 * ngb * m[i] / (4 * kPi * r)));
    int redo = !last && fabs(target - hi) > H_TOLERANCE * hi;
    if (redo) h[i] = gmax(hMin, gmax(0.5 * hi, gmin(2 * hi, target))); /* damped update */
    return redo;
}

/*
 * One hydro evaluation (sph.Solver.Forces) over the n bodies of a state.
 * kind 3 = gas; only alive gas takes part.  h and rho are read and updated
 * in place (h: smoothing lengths, rho: densities); u is the internal energy.
 * active (may be NULL = all) selects the evaluated bodies; step (may be
 * NULL) holds the bodies' current time steps for the limiter.  The outputs
 * ax..dt of evaluated bodies are written; wake (float64 bits) of neighbours
 * is lowered.  gamma, ngb, alpha of 0 take the defaults 5/3, 40, 1.
 * Returns the number of density evaluations (sum over iterations).
 */
int64_t or_sph_forces(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                      const double *vy, const double *vz, const double *m, const uint8_t *kind,
                      const uint8_t *alive, const double *u, double *h, double *rho, double gamma, double ngb,
                      double alpha, double softening, const uint8_t *active, const double *step, double *ax,
                      double *ay, double *az, double *dudt, double *divv, double *dt, uint64_t *wake) {
    int32_t *gas = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    int32_t ngas = 0;
    for (int64_t i = 0; i < n; i++)
        if (kind[i] == 3 && alive[i]) gas[ngas++] = (int32_t)i;
#define IS_ACTIVE(i) (active == NULL || active[i])
    if (ngas < 2) {
        for (int32_t k = 0; k < ngas; k++) {
            int32_t i = gas[k];
            if (IS_ACTIVE(i)) {
                ax[i] = ay[i] = az[i] = 0;
                dudt[i] = divv[i] = 0;
                dt[i] = INFINITY;
            }
        }
        free(gas);
        return 0;
    }
    if (gamma == 0) gamma = 5.0 / 3;
    if (ngb == 0) ngb = 40;
    if (alpha == 0) alpha = 1;
    /* initH, sph.go:277-302 */
    int missing = 0;
    for (int32_t k = 0; k < ngas; k++)
        if (h[gas[k]] <= 0) { missing = 1; break; }
    if (missing) {
        double minX = INFINITY, minY = INFINITY, minZ = INFINITY;
        double maxX = -INFINITY, maxY = -INFINITY, maxZ = -INFINITY;
        for (int32_t k = 0; k < ngas; k++) {
            int32_t i = gas[k];
            minX = gmin(minX, x[i]); maxX = gmax(maxX, x[i]);
            minY = gmin(minY, y[i]); maxY = gmax(maxY, y[i]);
            minZ = gmin(minZ, z[i]); maxZ = gmax(maxZ, z[i]);
        }
        double vol = gmax((maxX - minX) * (maxY - minY) * (maxZ - minZ), 1e-30);
        double guess = 0.5 * or_cbrt(3 * ngb * vol / (4 * kPi * (double)ngas));
        for (int32_t k = 0; k < ngas; k++)
            if (h[gas[k]] <= 0) h[gas[k]] = guess;
    }
    double hMin = H_MIN_SOFTENING * softening;
    sph_tree *t = (sph_tree *)or_sph_tree_new();
    or_sph_tree_build(t, gas, ngas, x, y, z, h);

    /* leaves holding evaluated bodies */
    int32_t *leaves = (int32_t *)malloc(sizeof(int32_t) * (t->nleaves ? t->nleaves : 1));
    int32_t nl = 0;
    for (int32_t l = 0; l < t->nleaves; l++) {
        const sph_node *nd = &t->nodes[t->leaves[l]];
        for (int32_t k = nd->lo; k < nd->hi; k++)
            if (IS_ACTIVE(t->order[k])) { leaves[nl++] = t->leaves[l]; break; }
    }

    uint8_t *again = (uint8_t *)calloc(ngas, 1);
    int32_t *todo = (int32_t *)malloc(sizeof(int32_t) * ngas);
    int32_t ntodo = 0;
    int32_t *cand = NULL, candCap = 0;
    int64_t evals = 0;
    for (int it = 0; it == 0 || ntodo > 0; it++) {
        int last = it == MAX_DENSITY_ITERATIONS - 1;
        if (it == 0) {
            for (int32_t l = 0; l < nl; l++) {
                int32_t nc = candidates(t, leaves[l], 0, &cand, &candCap);
                const sph_node *nd = &t->nodes[leaves[l]];
                for (int32_t k = nd->lo; k < nd->hi; k++) {
                    int32_t i = t->order[k];
                    if (!IS_ACTIVE(i)) { again[k] = 0; continue; }
                    density_ctx d = {m, h[i], 0.0};
                    scan(t, cand, nc, x[i], y[i], z[i], 2 * d.hi, 0, density_term, &d);
                    again[k] = set_h(i, d.rho, last, hMin, ngb, m, h, rho);
                    evals++;
                }
            }
            ntodo = 0;
            for (int32_t l = 0; l < nl; l++) {
                const sph_node *nd = &t->nodes[leaves[l]];
                for (int32_t k = nd->lo; k < nd->hi; k++)
                    if (again[k]) todo[ntodo++] = t->order[k];
            }
        } else {
            for (int32_t k = 0; k < ntodo; k++) {
                int32_t i = todo[k];
                density_ctx d = {m, h[i], 0.0};
                each(t, x[i], y[i], z[i], 2 * d.hi, 0, density_term, &d);
                again[k] = set_h(i, d.rho, last, hMin, ngb, m, h, rho);
                evals++;
            }
            int32_t j = 0;
            for (int32_t k = 0; k < ntodo; k++)
                if (again[k]) todo[j++] = todo[k];
            ntodo = j;
        }
        if (ntodo > 0) or_sph_tree_refit_h(t, h);
    }
    or_sph_tree_sort_leaves(t);

    double *P = (double *)malloc(sizeof(double) * n), *C = (double *)malloc(sizeof(double) * n);
    for (int32_t k = 0; k < ngas; k++) {
        int32_t i = gas[k];
        double ui = gmax(u[i], 0);
        P[i] = (gamma - 1) * rho[i] * ui;
        C[i] = sqrt(gamma * (gamma - 1) * ui);
    }
    for (int32_t l = 0; l < nl; l++) {
        int32_t nc = candidates(t, leaves[l], 1, &cand, &candCap);
        const sph_node *nd = &t->nodes[leaves[l]];
        for (int32_t k = nd->lo; k < nd->hi; k++) {
            int32_t i = t->order[k];
            if (!IS_ACTIVE(i)) continue;
            force_ctx f;
            memset(&f, 0, sizeof f);
            f.x = x; f.y = y; f.z = z; f.vx = vx; f.vy = vy; f.vz = vz; f.m = m; f.h = h; f.rho = rho;
            f.P = P; f.C = C; f.step = step; f.i = i; f.alpha = alpha;
            f.hi = h[i]; f.rhoi = rho[i];
            f.fi = P[i] / (f.rhoi * f.rhoi);
            f.vsig = 2 * C[i];
            f.minStep = INFINITY;
            f.maxStep = 0.0;
            scan(t, cand, nc, x[i], y[i], z[i], 2 * f.hi, 1, force_term, &f);
            ax[i] = f.ax; ay[i] = f.ay; az[i] = f.az;
            dudt[i] = f.du;
            divv[i] = f.dv / f.rhoi;
            double d = LIMITER_RATIO * f.minStep;
            if (f.vsig > 0) d = gmin(d, COURANT * f.hi / f.vsig);
            dt[i] = d;
            if (f.maxStep > LIMITER_RATIO * d) {
                wake_ctx w = {step, wake, d};
                scan(t, cand, nc, x[i], y[i], z[i], 2 * f.hi, 1, wake_term, &w);
            }
        }
    }
#undef IS_ACTIVE
    free(P); free(C); free(cand); free(again); free(todo); free(leaves); free(gas);
    or_sph_tree_free(t);
    return evals;
}
