"""Generates the golden fixtures of the FP64 oracle used by the GPU parity tests.

    python tests/golden/make_golden.py cavity2d      # 60^2 cells x 41^2 velocities, 2000 steps
    python tests/golden/make_golden.py channel       # 200x40 cells x 32^2 velocities, 2000 steps
    python tests/golden/make_golden.py cavity3d      # 48^3 cells x 24^3 velocities, 20 steps
    python tests/golden/make_golden.py delaunay      # ~200k triangles x 64^2 velocities, 5 steps

Each fixture stores rho, U, T, q at a few step counts and the total-mass
history, so a GPU run can be compared without re-running the CPU solver.

    python tests/golden/make_golden.py channel --floors

adds to an existing fixture, for its snapshots, the round-off floor of
each field: the relative L2 distance between the FP64 oracle and two equally
valid evaluations of the same arithmetic -- the oracle built with
extended-precision intermediates and velocity-space sums (floor_ext_*) and
with every velocity pass summed in reverse order (floor_rev_*).  Where a floor exceeds the parity bar,
the bar cannot be resolved in FP64 arithmetic (fields that are sums cancelling
almost to zero, e.g. the velocity one step after rest, or the heat flux of
cells along two walls).  The extended-precision oracle's fields themselves are
stored too (ext_*): a GPU result that misses 1e-12 against the FP64 oracle must
still be within 1e-12 of them, which shows that the gap is the FP64 oracle's
own round-off and not a difference of the step.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dugksfoam_amd import cases  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

PLANS = {
    "cavity2d": dict(case=lambda: cases.config("cavity2d"), snaps=(1, 10, 100, 500, 1000, 2000), cache=True,
                     floor_until=2000),
    "channel": dict(case=lambda: cases.config("channel"), snaps=(1, 10, 100, 500, 1000, 2000), cache=True,
                    floor_until=2000),
    # 3-D cavity (48^3 x 24^3) to step 20: past the first steps every cell's fields differ in the
    # low bits, so the fixture keeps every SAMPLE-th cell (7 is prime to 48: all planes)
    "cavity3d": dict(case=lambda: cases.config("cavity3d"), snaps=(1, 2, 5, 10, 20), cache=False,
                     floor_until=20, sample=7),
    # Delaunay cavity at full size (~200k triangles x 64^2 velocities): its fields do not compress
    # (every cell differs), so the fixture keeps every SAMPLE-th cell of them
    "delaunay": dict(case=lambda: cases.config("delaunay"), snaps=(1, 2, 5), cache=False, floor_until=5,
                     sample=7),
    # the 256-case batch in FP64 (its own dt and muRef per case)
    "batch": dict(case=lambda: cases.config("batch"), snaps=(1, 2), cache=False, floor_until=2),
}


def _sampled(plan, v):
    st = plan.get("sample", 1)
    return v[::st] if st > 1 else v


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def floors(name: str, threads: int, variant: str):
    plan = PLANS[name]
    path = os.path.join(HERE, f"{name}.npz")
    out = dict(np.load(path))
    c = plan["case"]()
    o = Oracle(c, threads=threads, cache_faces=plan["cache"], variant=variant)
    o.init_equilibrium()
    n = 0
    for s in plan["snaps"]:
        if s > plan["floor_until"]:
            break
        while n < s:
            o.step(nsteps=1)
            n += 1
            if n % 50 == 0 or s <= 20:
                print(f"{name} {variant}: step {n}", flush=True)
        m = o.macro()
        for k, v in m.items():
            v = _sampled(plan, v)
            out[f"floor_{variant}_{k}_{s}"] = np.array(rel_l2(v, out[f"{k}_{s}"]))
            if variant == "ext":
                out[f"ext_{k}_{s}"] = v
        print(name, variant, "step", s, {k: float(out[f"floor_{variant}_{k}_{s}"]) for k in m}, flush=True)
    np.savez_compressed(path, **out)
    print("updated", path)


def run(name: str, threads: int):
    plan = PLANS[name]
    c = plan["case"]()
    o = Oracle(c, threads=threads, cache_faces=plan["cache"])
    o.init_equilibrium()
    out = {}
    mass = [o.total_mass()]
    t0 = time.time()
    n = 0
    for s in plan["snaps"]:
        while n < s:
            o.step(nsteps=1)
            n += 1
            mass.append(o.total_mass())
            if n % 50 == 0 or s <= 20:
                print(f"{name}: step {n} ({time.time() - t0:.0f} s)", flush=True)
        m = o.macro()
        for k, v in m.items():
            out[f"{k}_{s}"] = _sampled(plan, v)
    out["mass"] = np.array(mass)
    out["snaps"] = np.array(plan["snaps"])
    out["dt"] = np.array(c.dt)
    out["sample"] = np.array(plan.get("sample", 1))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    nthreads = int(args[1]) if len(args) > 1 else os.cpu_count()
    if "--floors" in sys.argv:
        for variant in ("ext", "rev"):
            floors(args[0], nthreads, variant)
    elif "--ext" in sys.argv:
        floors(args[0], nthreads, "ext")
    else:
        run(args[0], nthreads)
