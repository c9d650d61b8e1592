"""Seeded initial conditions for the force-step workloads.

The formulas follow the reference's scene builders; the random streams are
numpy's PCG64 (the reference uses Go's math/rand/v2, which cannot be
reproduced here), so the inputs have the same distributions, not the same
bits.  Everything is vectorised so that 16M-particle inputs build in seconds.

  plummer          internal/scene/scene.go:378-406 (radius, speed rejection,
                   isotropic directions), randomDirection :482-487
  hernquist_halo   internal/scene/scene.go:411-456
  toomre_disk      internal/scene/scene.go:111-117 (Gamma(2, R_d) radii),
                   :140-238 (Toomre-Q dispersions, asymmetric drift,
                   sech^2 layer z = z0 atanh(2u-1))
  remove_bulk_motion internal/scene/scene.go:490-504

A ``State`` is a dict of contiguous arrays: px, py, pz, vx, vy, vz, m
(float64) and kind (uint8: 0 star, 1 black hole, 2 dark matter, 3 gas).
"""
from __future__ import annotations

import math

import numpy as np

KIND_STAR, KIND_HOLE, KIND_DARK, KIND_GAS = 0, 1, 2, 3


def _random_direction(rng, n):
    z = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * math.pi * rng.random(n)
    r = np.sqrt(1.0 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def _state(pos, vel, m, kind):
    n = len(m)
    return {
        "px": np.ascontiguousarray(pos[:, 0]), "py": np.ascontiguousarray(pos[:, 1]),
        "pz": np.ascontiguousarray(pos[:, 2]), "vx": np.ascontiguousarray(vel[:, 0]),
        "vy": np.ascontiguousarray(vel[:, 1]), "vz": np.ascontiguousarray(vel[:, 2]),
        "m": np.ascontiguousarray(m, dtype=np.float64),
        "kind": np.full(n, kind, np.uint8) if np.isscalar(kind) else np.ascontiguousarray(kind, np.uint8),
    }


def concat(*states):
    return {k: np.concatenate([s[k] for s in states]) for k in states[0]}


def plummer(n, seed, mass=1.0, a=1.0, max_frac=0.999, kind=KIND_STAR):
    """Equal-mass Plummer sphere in isotropic equilibrium (G = 1)."""
    rng = np.random.default_rng([seed, 0x91])
    r = a / np.sqrt((rng.random(n) * max_frac + 1e-6) ** (-2.0 / 3.0) - 1.0)
    q = np.empty(n)
    todo = np.arange(n)
    while len(todo):
        qq = rng.random(len(todo))
        ok = 0.1 * rng.random(len(todo)) < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    ve = math.sqrt(2.0 * mass) * (r * r + a * a) ** -0.25
    pos = _random_direction(rng, n) * r[:, None]
    vel = _random_direction(rng, n) * (q * ve)[:, None]
    return _state(pos, vel, np.full(n, mass / n), kind)


def hernquist_total(mass, a, rmax):
    u = rmax / (rmax + a)
    return mass / (u * u)


def hernquist_enclosed(mass, a, rmax, r):
    r = np.minimum(r, rmax)
    return hernquist_total(mass, a, rmax) * r * r / ((r + a) * (r + a))


def hernquist_sigma2(m, a, r):
    x = r / a
    return m / (12 * a) * (12 * x * (1 + x) ** 3 * np.log((1 + x) / x)
                           - x / (1 + x) * (25 + 52 * x + 42 * x * x + 12 * x ** 3))


def hernquist_halo(n, seed, mass=1.0, a=1.0, rmax=20.0, kind=KIND_DARK):
    """Hernquist sphere truncated at rmax, Jeans-dispersion Gaussian velocities
    capped below 0.95 v_esc."""
    rng = np.random.default_rng([seed, 0x4E57])
    mtot = hernquist_total(mass, a, rmax)
    smax = math.sqrt(mass / mtot)
    su = smax * np.sqrt(rng.random(n))
    r = a * su / (1 - su)
    sigma = np.sqrt(np.maximum(0.0, hernquist_sigma2(mtot, a, r)))
    vesc = np.sqrt(2 * mtot / (r + a))
    vel = np.empty((n, 3))
    todo = np.arange(n)
    while len(todo):
        v = rng.standard_normal((len(todo), 3)) * sigma[todo, None]
        ok = np.linalg.norm(v, axis=1) < 0.95 * vesc[todo]
        vel[todo[ok]] = v[ok]
        todo = todo[~ok]
    pos = _random_direction(rng, n) * r[:, None]
    return _state(pos, vel, np.full(n, mass / n), kind)


def toomre_disk(n, seed, mass=1.0, rd=1.0, z0=0.1, Q=1.5, rmin=0.0, rmax=10.0, eps=0.02,
                halo_mass=0.0, halo_a=5.0, halo_rmax=50.0, kind=KIND_STAR):
    """Equal-mass exponential disk with Toomre-Q dispersions."""
    rng = np.random.default_rng([seed, 0x9A1A])
    radii = np.empty(n)
    todo = np.arange(n)
    while len(todo):
        r = -rd * np.log(rng.random(len(todo)) * rng.random(len(todo)))
        ok = (r >= rmin) & (r <= rmax)
        radii[todo[ok]] = r[ok]
        todo = todo[~ok]
    radii.sort()
    masses = np.full(n, mass / n)
    cum = np.cumsum(masses)
    eps2 = eps * eps

    def enclosed(r):
        k = np.searchsorted(radii, r)
        m = np.where(k > 0, cum[np.maximum(k - 1, 0)], 0.0)
        if halo_mass > 0:
            m = m + hernquist_enclosed(halo_mass, halo_a, halo_rmax, r)
        return m

    def omega2(r):
        r2 = r * r + eps2
        return enclosed(r) / (r2 * np.sqrt(r2))

    gamma_cdf = lambda r: 1 - np.exp(-r / rd) * (1 + r / rd)
    norm = gamma_cdf(rmax) - gamma_cdf(rmin)
    r = radii
    w2 = omega2(r)
    h = 0.01 * r
    kappa2 = np.maximum(r * (omega2(r + h) - omega2(r - h)) / (2 * h) + 4 * w2, 1e-12)
    kappa, omega = np.sqrt(kappa2), np.sqrt(w2)
    vc = omega * r
    sigma = mass * np.exp(-r / rd) / (2 * math.pi * rd * rd * norm)
    sig_r = np.minimum(Q * 3.36 * sigma / kappa, 0.6 * vc)
    sig_phi = sig_r * kappa / (2 * omega)
    sig_z = np.sqrt(math.pi * sigma * z0)
    vphi = np.sqrt(np.maximum(0.0, vc * vc + sig_r * sig_r * (1 - kappa2 / (4 * w2) - 2 * r / rd)))
    phi = rng.random(n) * 2 * math.pi
    z = z0 * np.arctanh(2 * rng.random(n) - 1)
    sinp, cosp = np.sin(phi), np.cos(phi)
    vr = rng.standard_normal(n) * sig_r
    vt = vphi + rng.standard_normal(n) * sig_phi
    pos = np.stack([r * cosp, r * sinp, z], axis=1)
    vel = np.stack([vr * cosp - vt * sinp, vr * sinp + vt * cosp, rng.standard_normal(n) * sig_z], axis=1)
    return _state(pos, vel, masses, kind)


def remove_bulk_motion(s):
    m = s["m"]
    mt = m.sum()
    if mt == 0:
        return s
    for p, v in (("px", "vx"), ("py", "vy"), ("pz", "vz")):
        s[p] -= (m * s[p]).sum() / mt
        s[v] -= (m * s[v]).sum() / mt
    return s


def _rotate_x(s, ang):
    c, sn = math.cos(ang), math.sin(ang)
    for p, q in (("py", "pz"), ("vy", "vz")):
        a, b = s[p].copy(), s[q].copy()
        s[p], s[q] = c * a - sn * b, sn * a + c * b
    return s


def _shift(s, pos, vel):
    for k, d in zip(("px", "py", "pz"), pos):
        s[k] += d
    for k, d in zip(("vx", "vy", "vz"), vel):
        s[k] += d
    return s


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
PLUMMER_16K = dict(name="plummer-16k", n=16384, theta=0.5, eps=0.01, dt=1e-3)
DISK_HALO_1M = dict(name="disk-halo-1m", n=1048576, thetas=(0.5, 0.7), eps=0.02)
TWO_DISK_4M = dict(name="two-disk-4m", n=4194304, theta=0.5, eps=0.02, steps=100, rungs=4)
BATCHED_256 = dict(name="batched-256", systems=256, nmin=1024, nmax=65536, theta=0.6)
CLUMPS_16M = dict(name="clumps-16m", n=16777216, clumps=1000, theta=0.5, eps=1e-4)


def plummer_equal_mass(n=16384, seed=1):
    """Equal-mass Plummer sphere, G = M = a = 1, centre of mass at rest."""
    return remove_bulk_motion(plummer(n, seed, 1.0, 1.0))


def disk_halo(n=1048576, seed=2, disk_frac=0.7, halo_ratio=5.0, eps=0.02):
    """70 % exponential disk (R_d 1, z0 0.1, Q 1.5, truncated at 10 R_d) in a
    Hernquist halo (a 5, truncated at 50, mass 5x the disk)."""
    nd = int(round(disk_frac * n))
    nh = n - nd
    disk = toomre_disk(nd, seed, mass=1.0, rd=1.0, z0=0.1, Q=1.5, rmin=0.0, rmax=10.0, eps=eps,
                       halo_mass=halo_ratio, halo_a=5.0, halo_rmax=50.0, kind=KIND_STAR)
    halo = hernquist_halo(nh, seed + 1000, mass=halo_ratio, a=5.0, rmax=50.0, kind=KIND_DARK)
    return remove_bulk_motion(concat(disk, halo))


def two_disk_encounter(n=4194304, seed=3, pericentre=5.0, inclination_deg=30.0, start_sep=30.0):
    """Two copies of the disk+halo model on a parabolic orbit."""
    half = n // 2
    g1 = disk_halo(half, seed)
    g2 = _rotate_x(disk_halo(n - half, seed + 1), math.radians(inclination_deg))
    m1, m2 = g1["m"].sum(), g2["m"].sum()
    M = m1 + m2
    q = pericentre
    # parabola: r = 2q / (1 + cos f); h = sqrt(2 M q)
    cosf = 2 * q / start_sep - 1
    f = -math.acos(max(-1.0, min(1.0, cosf)))
    r = start_sep
    hmom = math.sqrt(2 * M * q)
    rel_pos = np.array([r * math.cos(f), r * math.sin(f), 0.0])
    vr, vt = M / hmom * math.sin(f), hmom / r
    rel_vel = np.array([vr * math.cos(f) - vt * math.sin(f), vr * math.sin(f) + vt * math.cos(f), 0.0])
    _shift(g1, -rel_pos * m2 / M, -rel_vel * m2 / M)
    _shift(g2, rel_pos * m1 / M, rel_vel * m1 / M)
    return concat(g1, g2)


def batched_sizes(systems=256, nmin=1024, nmax=65536, seed=4):
    rng = np.random.default_rng([seed, 0xB7])
    return np.floor(np.exp(rng.uniform(math.log(nmin), math.log(nmax), systems))).astype(np.int64)


def batched_systems(systems=256, nmin=1024, nmax=65536, seed=4):
    """Independent systems: even index Plummer, odd index Hernquist; each
    with theta 0.6 and eps = 0.01 N^(-1/3)."""
    sizes = batched_sizes(systems, nmin, nmax, seed)
    out = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            s = plummer(int(n), seed * 1000 + i, 1.0, 1.0)
        else:
            s = hernquist_halo(int(n), seed * 1000 + i, 1.0, 1.0, 20.0, kind=KIND_STAR)
        out.append(remove_bulk_motion(s))
    eps = [0.01 * float(n) ** (-1.0 / 3.0) for n in sizes]
    return out, eps


def clumps(n=16777216, nclumps=1000, seed=5, scale=0.005, sigma=1.0):
    """Plummer clumps with log-normal masses (total 1) and uniform centres in
    the unit cube; particle counts are proportional to clump mass, so all
    particles have the same mass."""
    rng = np.random.default_rng([seed, 0xC1])
    centres = rng.random((nclumps, 3))
    cm = rng.lognormal(0.0, sigma, nclumps)
    cm /= cm.sum()
    counts = np.floor(cm * n).astype(np.int64)
    rem = n - counts.sum()
    order = np.argsort(-(cm * n - counts))
    counts[order[:rem]] += 1
    parts = []
    for c in range(nclumps):
        k = int(counts[c])
        if k == 0:
            continue
        s = plummer(k, seed * 100000 + c, float(cm[c]), scale)
        _shift(s, centres[c], (0.0, 0.0, 0.0))
        parts.append(s)
    s = concat(*parts)
    s["m"][:] = 1.0 / n
    return s


def sample_targets(n, count=4096, seed=7):
    """Seeded sample of target indices for the direct-sum error check."""
    rng = np.random.default_rng([seed, 0x5A])
    return np.sort(rng.choice(n, size=min(count, n), replace=False)).astype(np.int64)


def rel_errors(got, want):
    """Per-particle |a - a_ref| / |a_ref| (bh_test.go:77-86)."""
    d = np.linalg.norm(np.asarray(got) - np.asarray(want), axis=1)
    w = np.linalg.norm(np.asarray(want), axis=1)
    return d / w


def error_stats(got, want):
    e = np.sort(rel_errors(got, want))
    if len(e) == 0:
        return dict(p50=0.0, p99=0.0, max=0.0)
    return dict(p50=float(e[len(e) // 2]), p99=float(e[min(len(e) - 1, int(0.99 * len(e)))]),
                max=float(e[-1]))


# -- gas (hydro workloads) -------------------------------------------------------------


def _gas(s, u):
    n = len(s["px"])
    s["kind"] = np.full(n, KIND_GAS, np.uint8)
    s["alive"] = np.ones(n, np.uint8)
    s["u"] = np.ascontiguousarray(u, np.float64)
    s["h"] = np.zeros(n)  # 0: the hydro makes a first guess (internal/sph/sph.go:277-302)
    s["rho"] = np.zeros(n)
    return s


def gas_cloud(n, seed):
    """Gaussian gas cloud of the hydro tests (internal/sph/sph_test.go:32-45):
    positions N(0, 1) with z halved, velocities N(0, 0.3), masses
    (0.5 + U) / n, internal energy 0.1 + 0.2 U."""
    rng = np.random.default_rng([seed, 0x6A5])
    pos = rng.standard_normal((n, 3)) * np.array([1.0, 1.0, 0.5])
    vel = rng.standard_normal((n, 3)) * 0.3
    m = (0.5 + rng.random(n)) / n
    u = 0.1 + 0.2 * rng.random(n)
    return _gas(_state(pos, vel, m, KIND_GAS), u)


def gas_sphere(n, seed=9, neighbors=40.0, h_from_density=False):
    """Plummer sphere of gas made hot enough to be roughly pressure supported
    (internal/sph/system_test.go:14-23): u = 0.2, velocities scaled by 0.3.
    With h_from_density the smoothing lengths start at the hydro's target for
    the sphere's analytic density, 0.5 (3 N m / 4 pi rho(r))^(1/3) (sph.go:117-
    118), instead of 0 (the hydro's box-volume guess)."""
    s = plummer(n, seed)
    for k in ("vx", "vy", "vz"):
        s[k] = s[k] * 0.3
    s = _gas(s, np.full(n, 0.2))
    if h_from_density:
        r2 = s["px"] ** 2 + s["py"] ** 2 + s["pz"] ** 2
        rho = 3.0 / (4 * math.pi) * (1 + r2) ** -2.5  # M = a = 1
        s["h"] = 0.5 * np.cbrt(3 * neighbors * s["m"] / (4 * math.pi * rho))
    return s


def gas_lattice(k=16, spacing=0.1):
    """k^3 bodies of unit mass on a cubic lattice, u = 1 (internal/sph/sph_test.go:133-143)."""
    a = np.arange(k) * spacing
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    n = len(pos)
    return _gas(_state(pos, np.zeros((n, 3)), np.ones(n), KIND_GAS), np.ones(n))
