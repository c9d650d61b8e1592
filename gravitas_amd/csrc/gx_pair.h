// gx_pair.h -- pair model of bodies and black holes (internal/sim/pair.go,
// relativity.go), FP64, shared by the potential and the split step.
#pragma once
#include "gx_common.h"

namespace gx {

// math.Max of Go for the finite, non-negative arguments used here.
__device__ __forceinline__ double go_max_d(double a, double b) { return a > b ? a : b; }

// SchwarzschildRadius, internal/sim/relativity.go:49-54 (0 when c <= 0).
__device__ __forceinline__ double schwarzschild(double m, double c) { return c <= 0 ? 0 : 2 * m / (c * c); }

// PairRS, pair.go:13-22: the hole's radius for a body-hole pair, else 0.
__device__ __forceinline__ double pair_rs(int ki, double mi, int kj, double mj, double c) {
    if (c <= 0 || (ki == kHole) == (kj == kHole)) return 0;
    return ki == kHole ? schwarzschild(mi, c) : schwarzschild(mj, c);
}

// PairForce, pair.go:26-34: f with acceleration f*d from a source of mass m.
__device__ __forceinline__ double pair_force(double m, double r2, double eps2, double rs) {
    const double re2 = r2 + eps2;
    const double re = sqrt(re2);
    if (rs == 0) return m / (re2 * re);
    const double den = go_max_d(re - rs, rs / 2);
    return m / (den * den * re);
}

// PairPotential, pair.go:37-47.
__device__ __forceinline__ double pair_potential(double mi, double mj, double r2, double eps2, double rs) {
    const double re = sqrt(r2 + eps2);
    if (rs == 0) return -mi * mj / re;
    if (re >= 1.5 * rs) return -mi * mj / (re - rs);
    return mi * mj * (-2 / rs + 4 / (rs * rs) * (re - 1.5 * rs));
}

// gamma, relativity.go:14-19: Lorentz factor of the proper velocity u.
__device__ __forceinline__ double gamma_of(double ux, double uy, double uz, double c) {
    if (c <= 0) return 1;
    return sqrt(1 + (ux * ux + uy * uy + uz * uz) / (c * c));
}

}  // namespace gx
