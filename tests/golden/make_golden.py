"""Generates the golden fixtures of the FP64 oracle used by the GPU parity tests.

    python tests/golden/make_golden.py cavity2d      # 60^2 cells x 41^2 velocities, 2000 steps
    python tests/golden/make_golden.py channel       # 200x40 cells x 32^2 velocities, 2000 steps
    python tests/golden/make_golden.py cavity3d      # 48^3 cells x 24^3 velocities, 2 steps

Each fixture stores rho, U, T, q at a few step counts and the total-mass
history, so a GPU run can be compared without re-running the CPU solver.
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
    "cavity2d": dict(case=lambda: cases.config("cavity2d"), snaps=(1, 10, 100, 500, 1000, 2000), cache=True),
    "channel": dict(case=lambda: cases.config("channel"), snaps=(1, 10, 100, 500, 1000, 2000), cache=True),
    "cavity3d": dict(case=lambda: cases.config("cavity3d"), snaps=(1, 2), cache=False),
}


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
            if n % 50 == 0:
                print(f"{name}: step {n} ({time.time() - t0:.0f} s)", flush=True)
        m = o.macro()
        for k, v in m.items():
            out[f"{k}_{s}"] = v
    out["mass"] = np.array(mass)
    out["snaps"] = np.array(plan["snaps"])
    out["dt"] = np.array(c.dt)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    run(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else os.cpu_count())
