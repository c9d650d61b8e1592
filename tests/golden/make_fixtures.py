"""Writes the golden fixtures of the force step: small seeded inputs and the
oracle's outputs on them, as .npz files (no pickles) next to this script.

The reference holds no golden vectors for this path (its bh tests are
invariants on self-generated inputs, bh_test.go), and its inputs come from Go's
math/rand, which cannot be reproduced here; these files pin both the inputs and
the expected outputs, so the oracle and the HIP path are compared on the same
bytes on every machine, whatever numpy's generators do later.

    python tests/golden/make_fixtures.py        # from the repository root

Files:
  plummer_512.npz  equal-mass Plummer sphere (scene.go:389-405 distribution):
                   Morton keys and permutation (build.go:19-32), accelerations at
                   theta 0.5, at theta 0 (= 0.9, bh.go:69-74) and monopole-only,
                   the potential energy (potential.go:12-42) and one KDK step
                   (sim.go:262-281).
  mixed_768.npz    Plummer stars + Hernquist dark halo + 2 holes + 3 dead
                   bodies: accelerations with dark matter on and off
                   (sim.go:93-98), potential energy with c = 0 and c = 40
                   (pair terms of the holes), one split step with holes
                   (split.go:164-199).
  encounter_768.npz two-disk encounter on 4 rungs: one global block step
                   (blockstep.go:130-198): rung populations, events, state.
  gas_768.npz      SPH hydro (sph.go:83-263): a gas cloud with stars and dead
                   bodies, h guessed from the box, a step array; h, rho,
                   forces, du/dt, div v, dt, wake and the density-sum count.
  holes_512.npz    Plummer cluster with two moving holes, one spinning, c = 4:
                   25 split steps with frame dragging and retarded+v hole
                   gravity (split.go, framedrag.go, retarded.go): substeps,
                   final state and spins, the holes' sources for 16 observers.
  edges.npz        two stars and a hole (bh_test.go:137-147), 100 coincident
                   stars and one more (bh_test.go:149-161), a leaf-size
                   boundary of 24 and 25 stars (build.go:134-160), one star.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gravitas_amd import ics  # noqa: E402

STATE = ("px", "py", "pz", "vx", "vy", "vz", "m", "kind", "alive")


def inputs(s):
    n = len(s["px"])
    out = {k: np.ascontiguousarray(s[k]) for k in STATE if k in s}
    out.setdefault("alive", np.ones(n, np.uint8))
    return out


def acc(o, s, eps, dm_off=False):
    n = len(s["px"])
    return o.accelerations(s["px"], s["py"], s["pz"], s["m"], s["kind"], s["alive"], eps=eps, dm_off=dm_off,
                           out=tuple(np.zeros(n) for _ in range(3)))


def plummer_512():
    s = inputs(ics.plummer(512, 101))
    eps, dt = 0.01, 1e-3
    o = oracle.Solver(theta=0.5)
    f = {"theta": 0.5, "eps": eps, "dt": dt}
    f["acc"] = acc(o, s, eps)
    f["keys"], f["order"], f["n_nodes"] = o.keys(), o.order(), o.n_nodes
    f["acc_theta0"] = acc(oracle.Solver(theta=0.0), s, eps)
    f["acc_mono"] = acc(oracle.Solver(theta=0.5, monopole_only=True), s, eps)
    f["potential"] = o.potential_energy(s["px"], s["py"], s["pz"], s["m"], s["kind"], s["alive"], eps=eps)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    st["acc"] = f["acc"].copy()
    oracle.kick_drift_kick(st, lambda x, y, z: acc(o, dict(s, px=x, py=y, pz=z), eps), dt)
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        f["kdk_" + k] = st[k]
    f["kdk_acc"] = st["acc"]
    return s, f


def mixed_768():
    s = ics.concat(ics.plummer(384, 102), ics.hernquist_halo(384, 103, mass=2.0, a=2.0, rmax=20.0))
    s["kind"][[3, 300]] = ics.KIND_HOLE
    s["m"][[3, 300]] = [0.2, 0.1]
    s["px"][3], s["py"][3], s["pz"][3] = 0.02, -0.01, 0.03  # a hole in the core: stars near it substep
    s["alive"] = np.ones(768, np.uint8)
    s["alive"][[5, 99, 500]] = 0
    s = inputs(s)
    eps, theta, dt = 0.01, 0.6, 2e-2
    o = oracle.Solver(theta=theta)
    f = {"theta": theta, "eps": eps, "dt": dt}
    f["acc"] = acc(o, s, eps)
    f["acc_dmoff"] = acc(o, s, eps, dm_off=True)
    for c in (0.0, 40.0):
        f[f"potential_c{int(c)}"] = o.potential_energy(s["px"], s["py"], s["pz"], s["m"], s["kind"], s["alive"],
                                                       eps=eps, c=c)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    f["split_substeps"] = oracle.split_step(o, st, s["m"], s["kind"], s["alive"], eps, dt, acc_valid=False)
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        f["split_" + k] = st[k]
    return s, f


def encounter_768():
    s = ics.two_disk_encounter(768, seed=104)
    s["alive"] = np.ones(768, np.uint8)
    s["alive"][[3, 600]] = 0
    s["kind"][7] = ics.KIND_HOLE
    s = inputs(s)
    eps, theta, eta, levels = 0.02, 0.5, 0.025, 4
    o = oracle.Solver(theta=theta)
    fn = lambda x, y, z: acc(o, dict(s, px=x, py=y, pz=z), eps)
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz")}
    st["acc"] = fn(st["px"], st["py"], st["pz"])
    a = np.linalg.norm(st["acc"], axis=1)
    a = a[a > 0]
    # global step that puts about 5 % of the stars on the top rung
    T = 4 * float(np.quantile(np.sqrt(2 * eta * eps / a), 0.05))
    r = oracle.block_step(st, fn, T, levels, eta, eps, s["kind"], s["alive"])
    f = {"theta": theta, "eps": eps, "eta": eta, "levels": levels, "T": T,
         "rung_count": np.array(r["rung_count"], np.int64), "events": np.array(r["events"], np.int64).reshape(-1, 2)}
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        f["block_" + k] = st[k]
    f["block_acc"] = st["acc"]
    return s, f


def edges():
    """Small cases at the default theta (0 = 0.9), each with its own eps."""
    f = {}
    cases = {}
    # bh_test.go:137-147: stars 1@x=1, 2@x=2 and a hole of mass 4 at the origin, eps 0
    cases["two"] = {"px": np.array([1.0, 2.0, 0.0]), "m": np.array([1.0, 2.0, 4.0]),
                    "kind": np.array([0, 0, ics.KIND_HOLE], np.uint8), "eps": 0.0}
    # bh_test.go:149-161: 100 stars at (1, 2, 3) and one at the origin, eps 0.01
    cases["coincident"] = {"px": np.r_[np.full(100, 1.0), 0.0], "py": np.r_[np.full(100, 2.0), 0.0],
                           "pz": np.r_[np.full(100, 3.0), 0.0], "m": np.ones(101), "eps": 0.01}
    rng = np.random.default_rng(105)
    for n in (24, 25):  # one leaf, and the first split (leaf size 24, build.go:134-160)
        cases[f"leaf{n}"] = {"px": rng.random(n), "py": rng.random(n), "pz": rng.random(n),
                             "m": rng.random(n) + 0.5, "eps": 0.01}
    cases["single"] = {"px": np.array([0.3]), "m": np.array([2.0]), "eps": 0.01}
    for name, c in cases.items():
        n = len(c["px"])
        s = {"px": c["px"], "py": c.get("py", np.zeros(n)), "pz": c.get("pz", np.zeros(n)), "m": c["m"],
             "kind": c.get("kind", np.zeros(n, np.uint8)), "alive": np.ones(n, np.uint8)}
        for k in ("px", "py", "pz", "m", "kind", "alive"):
            f[f"{name}__{k}"] = np.ascontiguousarray(s[k])
        f[f"{name}__eps"] = c["eps"]
        f[f"{name}__acc"] = acc(oracle.Solver(theta=0.0), s, c["eps"])
    return f


def gas_768():
    """A gas cloud with stars and dead bodies mixed in, h from the box guess
    (several density passes), a step array for the limiter: one
    sph.Solver.Forces evaluation (sph.go:83-263)."""
    s = ics.gas_cloud(768, 106)
    s["kind"][::13] = ics.KIND_STAR
    s["alive"][[7, 300]] = 0
    rng = np.random.default_rng(107)
    step = np.where(rng.random(768) < 0.1, 5.0, 1e-3)
    so = {k: v.copy() for k, v in s.items()}
    out, evals = oracle.sph_forces(so, softening=0.02, step=step)
    f = {"eps": 0.02, "step": step, "evals": evals, "h_out": so["h"], "rho_out": so["rho"]}
    for k in ("ax", "ay", "az", "dudt", "divv", "dt", "wake"):
        f[k] = out[k]
    return {k: s[k] for k in ("px", "py", "pz", "vx", "vy", "vz", "m", "kind", "alive", "u", "h", "rho")}, f


def holes_512():
    """A Plummer cluster with two moving holes, the first spinning, c = 4:
    25 split steps (split.go:164-199) with frame dragging (framedrag.go) and
    extrapolated retarded hole gravity (retarded.go, "retarded+v"), the time
    advancing by dt after each; the substeps of every step, the final state
    and spins, and where each hole is seen from 16 observers at the end."""
    s = ics.plummer(512, 108)
    s["kind"][[0, 1]] = ics.KIND_HOLE
    s["m"][[0, 1]] = [0.05, 0.03]
    s["px"][[0, 1]], s["py"][[0, 1]], s["pz"][[0, 1]] = [0.05, -0.1], [0.0, 0.05], [0.0, 0.0]  # in the core
    s["vx"][[0, 1]] = [0.8, -0.6]
    s["alive"] = np.ones(512, np.uint8)
    s["alive"][[9, 200]] = 0
    c, eps, theta, dt = 4.0, 0.01, 0.5, 1e-2
    for k in ("jx", "jy", "jz"):
        s[k] = np.zeros(512)
    s["jz"][0] = 0.5 * 0.05 ** 2 / c  # a* = 0.5: J = a* M^2 / c
    s = {**inputs(s), "jx": s["jx"], "jy": s["jy"], "jz": s["jz"]}
    o, r = oracle.Solver(theta=theta), oracle.Retard()
    st = {k: s[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz", "jx", "jy", "jz")}
    subs = []
    for k in range(25):
        subs.append(oracle.split_step(o, st, s["m"], s["kind"], s["alive"], eps, dt, c=c, acc_valid=k > 0, retard=r,
                                      time=k * dt, gravity="retarded+v"))
    obs = np.random.default_rng(109).uniform(-2, 2, (16, 3))
    f = {"theta": theta, "eps": eps, "dt": dt, "c": c, "steps": 25, "substeps": np.array(subs, np.int64),
         "obs": obs}
    for h in range(2):
        src = [r.source_pos(st, s["kind"], s["alive"], h, q, 25 * dt, dt, c, "retarded+v") for q in obs]
        f[f"source{h}"] = np.array([q if q is not None else [np.nan] * 3 for q in src])
    for k in st:
        f["out_" + k] = st[k]
    return s, f


def write(name, arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
    print(f"{name}: {os.path.getsize(path)} bytes, {len(arrays)} arrays")


def main():
    for name, make in (("plummer_512.npz", plummer_512), ("mixed_768.npz", mixed_768),
                       ("encounter_768.npz", encounter_768), ("gas_768.npz", gas_768), ("holes_512.npz", holes_512)):
        s, f = make()
        write(name, {**{"in_" + k: v for k, v in s.items()}, **f})
    write("edges.npz", edges())


if __name__ == "__main__":
    main()
