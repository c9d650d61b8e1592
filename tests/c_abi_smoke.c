/*
 * c_abi_smoke.c -- the backend's C ABI used from plain C, the way the Go
 * binding would (cgo calls the same functions with the same pointers).
 *
 * Without a usable GPU: gx_init must fail cleanly with a status code and a
 * message, and gx_solver_new must return NULL (exit 0, "no device").
 * With a GPU:
 *   - two stars, m = 1 at x = 1 and m = 2 at x = 2, eps = 0: the
 *     accelerations are (2, 0, 0) and (-1, 0, 0) (bh_test.go:137-147);
 *   - a hole gets 0 and a dead body keeps its value (bh.go:87-101);
 *   - gx_last_seconds is 0 before the first step and positive after one
 *     with timing off (LastGPUSeconds, gpu_darwin.go:155);
 *   - a KDK step on device arrays gives the same positions as one computed
 *     here from the step's own formulas (sim.go:302-343), using the forces
 *     returned by the host entry point;
 *   - argument errors come back as status codes, not crashes;
 *   - retarded hole gravity: after 200 split steps a hole moving at 0.3 c is
 *     seen on the observer's past light cone (retarded_test.go:22-38);
 *   - the SPH hydro from C: a 3x3x3 gas lattice, no net force on the centre
 *     body, densities and smoothing lengths positive (sph.go:83-263).
 * Prints "ok ..." and exits 0 on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gravitas_hip.h"

#include <hip/hip_runtime_api.h>

static int fail(const char *what) {
    fprintf(stderr, "FAIL: %s (%s)\n", what, gx_error());
    return 1;
}

int main(void) {
    char msg[256];
    int rc = gx_init(msg, sizeof msg);
    if (rc != GX_OK) {
        if (rc != GX_ERR_NO_DEVICE && rc != GX_ERR_CODE_OBJECT) return fail("unexpected gx_init status");
        if (gx_solver_new() != NULL) return fail("solver without a device");
        printf("no device: %s\n", msg);
        return 0;
    }
    char name[256];
    gx_name(name, sizeof name);
    gx_solver *s = gx_solver_new();
    if (!s) return fail("gx_solver_new");
    if (gx_solver_config(s, 0.5, 0, GX_WALK_FP64, 0) != GX_OK) return fail("config");
    if (gx_last_seconds(s) != 0) return fail("LastGPUSeconds must be 0 before the first step");

    /* two stars, a hole and a dead body */
    double x[4] = {1, 2, 5, 7}, y[4] = {0, 0, 0, 0}, z[4] = {0, 0, 0, 0}, m[4] = {1, 2, 3, 4};
    uint8_t kind[4] = {GX_KIND_STAR, GX_KIND_STAR, GX_KIND_HOLE, GX_KIND_STAR};
    uint8_t alive[4] = {1, 1, 1, 0};
    double ax[4] = {9, 9, 9, 42}, ay[4] = {9, 9, 9, 42}, az[4] = {9, 9, 9, 42};
    if (gx_accelerations(s, x, y, z, m, kind, alive, 4, 0.0, 0, ax, ay, az) != GX_OK) return fail("accelerations");
    if (ax[0] != 2 || ax[1] != -1 || ay[0] != 0 || az[1] != 0) return fail("two-star known answer");
    if (ax[2] != 0 || ay[2] != 0 || az[2] != 0) return fail("hole must get 0");
    if (ax[3] != 42 || ay[3] != 42 || az[3] != 42) return fail("dead body must be untouched");
    /* LastGPUSeconds (gpu_darwin.go:155) with timing off, as the Go binding configures it */
    if (!(gx_last_seconds(s) > 0)) return fail("LastGPUSeconds must be positive after a step");

    /* errors are status codes */
    if (gx_accelerations(s, NULL, y, z, m, kind, alive, 4, 0.0, 0, ax, ay, az) != GX_ERR_ARG)
        return fail("null array must be GX_ERR_ARG");
    if (gx_solver_config(s, -1.0, 0, GX_WALK_FP64, 0) != GX_ERR_ARG) return fail("negative theta must be GX_ERR_ARG");
    if (gx_solver_config(s, 0.5, 0, GX_WALK_FP64, 0) != GX_OK) return fail("config");

    /* a KDK step on device arrays (two stars only) */
    const int n = 2;
    const double dt = 0.01;
    double hx[2] = {1, 2}, hv[2] = {0.1, -0.2}, zero[2] = {0, 0}, mm[2] = {1, 2};
    uint8_t hk[2] = {0, 0}, ha[2] = {1, 1};
    double *d[10];
    for (int i = 0; i < 10; i++)
        if (hipMalloc((void **)&d[i], n * sizeof(double)) != hipSuccess) return fail("hipMalloc");
    uint8_t *dk, *da;
    if (hipMalloc((void **)&dk, n) != hipSuccess || hipMalloc((void **)&da, n) != hipSuccess) return fail("hipMalloc");
    /* d: x y z vx vy vz m ax ay az */
    hipMemcpy(d[0], hx, sizeof hx, hipMemcpyHostToDevice);
    hipMemcpy(d[1], zero, sizeof zero, hipMemcpyHostToDevice);
    hipMemcpy(d[2], zero, sizeof zero, hipMemcpyHostToDevice);
    hipMemcpy(d[3], hv, sizeof hv, hipMemcpyHostToDevice);
    hipMemcpy(d[4], zero, sizeof zero, hipMemcpyHostToDevice);
    hipMemcpy(d[5], zero, sizeof zero, hipMemcpyHostToDevice);
    hipMemcpy(d[6], mm, sizeof mm, hipMemcpyHostToDevice);
    hipMemcpy(dk, hk, sizeof hk, hipMemcpyHostToDevice);
    hipMemcpy(da, ha, sizeof ha, hipMemcpyHostToDevice);
    if (gx_step_dev(s, d[0], d[1], d[2], d[3], d[4], d[5], d[6], dk, da, n, 0.0, 0, dt, 0, d[7], d[8], d[9]) != GX_OK)
        return fail("gx_step_dev");
    double gx[2], gv[2];
    hipMemcpy(gx, d[0], sizeof gx, hipMemcpyDeviceToHost);
    hipMemcpy(gv, d[3], sizeof gv, hipMemcpyDeviceToHost);
    /* the same step on the host with forces from the host entry point */
    double px[2] = {1, 2}, vx[2] = {0.1, -0.2}, a[2], a0[2], a1[2];
    double py[2] = {0, 0}, pz[2] = {0, 0};
    if (gx_accelerations(s, px, py, pz, mm, hk, ha, 2, 0.0, 0, a, a0, a1) != GX_OK) return fail("accelerations");
    for (int i = 0; i < 2; i++) vx[i] += 0.5 * dt * a[i];
    for (int i = 0; i < 2; i++) px[i] += dt * vx[i];
    if (gx_accelerations(s, px, py, pz, mm, hk, ha, 2, 0.0, 0, a, a0, a1) != GX_OK) return fail("accelerations");
    for (int i = 0; i < 2; i++) vx[i] += 0.5 * dt * a[i];
    if (gx[0] != px[0] || gx[1] != px[1] || gv[0] != vx[0] || gv[1] != vx[1]) return fail("KDK step");
    for (int i = 0; i < 10; i++) hipFree(d[i]);
    hipFree(dk);
    hipFree(da);

    gx_stats st;
    if (gx_get_stats(s, &st) != GX_OK || st.n_stars != 2) return fail("stats");

    /* retarded hole gravity (retarded.go): a light hole moving along x at
     * 0.3 c (c = 10) and a star at x = 50, 200 split steps of dt 0.004 with
     * the holes' histories; the hole seen from (1, 4, -2) lies on that
     * point's past light cone, behind the hole's current position */
    {
        const double c = 10.0, dtr = 0.004;
        const double u = 3.0 / sqrt(1.0 - 0.09); /* ProperVelocity of v = 3 */
        double rx[2] = {0, 50}, rv[2] = {u, 0}, rz[2] = {0, 0}, rm[2] = {1e-12, 1e-12};
        uint8_t rk[2] = {GX_KIND_HOLE, GX_KIND_STAR}, ra[2] = {1, 1};
        double *e[10];
        uint8_t *ek, *ea;
        for (int i = 0; i < 10; i++)
            if (hipMalloc((void **)&e[i], 2 * sizeof(double)) != hipSuccess) return fail("hipMalloc");
        if (hipMalloc((void **)&ek, 2) != hipSuccess || hipMalloc((void **)&ea, 2) != hipSuccess) return fail("hipMalloc");
        hipMemcpy(e[0], rx, sizeof rx, hipMemcpyHostToDevice);
        hipMemcpy(e[3], rv, sizeof rv, hipMemcpyHostToDevice);
        hipMemcpy(e[6], rm, sizeof rm, hipMemcpyHostToDevice);
        for (int i = 1; i < 6; i++)
            if (i != 3) hipMemcpy(e[i], rz, sizeof rz, hipMemcpyHostToDevice);
        hipMemcpy(ek, rk, sizeof rk, hipMemcpyHostToDevice);
        hipMemcpy(ea, ra, sizeof ra, hipMemcpyHostToDevice);
        gx_retard *r = gx_retard_new();
        if (!r) return fail("gx_retard_new");
        double T = 0;
        for (int k = 0; k < 200; k++, T += dtr) {
            int32_t sub = 0;
            if (gx_split_step_retarded_dev(s, e[0], e[1], e[2], e[3], e[4], e[5], e[6], ek, ea, NULL, NULL, NULL, NULL, 2,
                                           0.01, 0, dtr, c, k > 0, e[7], e[8], e[9], r, T, GX_GRAVITY_RETARDED,
                                           &sub) != GX_OK)
                return fail("gx_split_step_retarded_dev");
        }
        if (gx_retard_count(r) != 1) return fail("one hole, one history");
        double ob[3] = {1, 4, -2}, *o, *sxyz;
        uint8_t *vis, hvis = 0;
        if (hipMalloc((void **)&o, 3 * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&sxyz, 3 * sizeof(double)) != hipSuccess || hipMalloc((void **)&vis, 1) != hipSuccess)
            return fail("hipMalloc");
        hipMemcpy(o, ob, sizeof ob, hipMemcpyHostToDevice);
        if (gx_retard_source_pos_dev(s, r, e[0], e[1], e[2], ek, ea, NULL, 2, 0, o, o + 1, o + 2, 1, T, dtr, c,
                                     GX_GRAVITY_RETARDED, sxyz, sxyz + 1, sxyz + 2, vis) != GX_OK)
            return fail("gx_retard_source_pos_dev");
        double src[3], cur;
        hipMemcpy(src, sxyz, sizeof src, hipMemcpyDeviceToHost);
        hipMemcpy(&hvis, vis, 1, hipMemcpyDeviceToHost);
        hipMemcpy(&cur, e[0], sizeof cur, hipMemcpyDeviceToHost);
        const double dist = sqrt((ob[0] - src[0]) * (ob[0] - src[0]) + (ob[1] - src[1]) * (ob[1] - src[1]) +
                                 (ob[2] - src[2]) * (ob[2] - src[2]));
        if (!hvis || fabs(dist - c * (T - src[0] / 3.0)) > 1e-6 || !(src[0] < cur)) return fail("light cone");
        if (gx_split_step_retarded_dev(s, e[0], e[1], e[2], e[3], e[4], e[5], e[6], ek, ea, NULL, NULL, NULL, NULL, 2,
                                       0.01, 0, dtr, c, 1, e[7], e[8], e[9], r, T, 7, NULL) != GX_ERR_ARG)
            return fail("unknown gravity mode must be GX_ERR_ARG");
        gx_retard_free(r);
        for (int i = 0; i < 10; i++) hipFree(e[i]);
        hipFree(ek);
        hipFree(ea);
        hipFree(o);
        hipFree(sxyz);
        hipFree(vis);
    }
    gx_solver_free(s);

    /* the hydro (sph.Solver.Forces) on host arrays: a 3x3x3 gas lattice of
     * unit masses, spacing 0.1, u = 1, h guessed; by symmetry the centre body
     * feels no net force (to round-off), and every body gets rho > 0, h > 0 */
    enum { NL = 27 };
    double lx[NL], ly[NL], lz[NL], lv[NL] = {0}, lm[NL], lu[NL], lh[NL] = {0}, lr[NL] = {0};
    double hax[NL], hay[NL], haz[NL], hdu[NL], hdv[NL], hdt[NL];
    uint64_t wake[NL];
    uint8_t lk[NL], la[NL];
    const double inf = 1.0 / 0.0;
    for (int i = 0; i < NL; i++) {
        lx[i] = 0.1 * (i % 3);
        ly[i] = 0.1 * (i / 3 % 3);
        lz[i] = 0.1 * (i / 9);
        lm[i] = 1;
        lu[i] = 1;
        lk[i] = GX_KIND_GAS;
        la[i] = 1;
        memcpy(&wake[i], &inf, sizeof inf);
    }
    gx_hydro *hy = gx_hydro_new();
    if (!hy) return fail("gx_hydro_new");
    gx_gas_params gas = {0, 0, 0};
    if (gx_hydro_forces(hy, lx, ly, lz, lv, lv, lv, lm, lk, la, lu, lh, lr, NL, &gas, 0.0, NULL, NULL, hax, hay,
                        haz, hdu, hdv, hdt, wake) != GX_OK)
        return fail("gx_hydro_forces");
    for (int i = 0; i < NL; i++)
        if (!(lr[i] > 0) || !(lh[i] > 0) || !(hdt[i] > 0)) return fail("hydro density, h or dt");
    const double c = fabs(hax[13]) + fabs(hay[13]) + fabs(haz[13]), edge = fabs(hax[12]);
    if (!(c <= 1e-9 * edge) || !(edge > 0)) return fail("hydro lattice symmetry");
    if (gx_hydro_forces(hy, NULL, ly, lz, lv, lv, lv, lm, lk, la, lu, lh, lr, NL, &gas, 0.0, NULL, NULL, hax, hay,
                        haz, hdu, hdv, hdt, wake) != GX_ERR_ARG)
        return fail("hydro null array must be GX_ERR_ARG");
    gx_hydro_free(hy);
    printf("ok: %s; two stars (%g, %g), KDK x = (%.15g, %.15g)\n", name, ax[0], ax[1], gx[0], gx[1]);
    return 0;
}
