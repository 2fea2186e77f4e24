"""GPU: dugksFoam case directories through the front-end and the application loop.

The reference's demo/cavity (from its array fixture: 3600 cells, 28^2 velocities, lid 50 m/s,
rarefied argon-like gas, adjustTimeStep at maxCo 0.8), its demo/testSymmetry/symmetry (the
left half with a symmetryMod plane) and a cyclic channel written in OpenFOAM format run through dugksfoam_amd.solver.Solver on the HIP engine and on the FP64
oracle.  Bar: conftest.parity_ok on rho, U, T, q -- relative L2 <= 1e-12 against the FP64
oracle, or, where the oracle's own FP64 arithmetic is that far from the exact result (floor: its
distance from the oracle with every velocity sum reversed and from the extended-precision
oracle, both run live), within 1e-12 of the extended-precision oracle or no farther from it than
the FP64 oracle is -- and the same time-step sequence to 1e-13 (1e-8 with FP32 storage, whose bar
is 1e-5).
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, parity_ok, rel_l2
from dugksfoam_amd import cases, foam, solver
from dugksfoam_amd.engine import Engine
import foam_box

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _OracleBackend:
    def __init__(self, case, variant="fp64"):
        from oracle.oracle import Oracle
        self.o = Oracle(case, threads=16, variant=variant)
        self.steps = []

    def init_equilibrium(self, rho, U, T):
        self.o.init_equilibrium(rho, U, T)

    def step(self, dt, n):
        self.steps.append(dt)
        self.o.step(dt=dt, nsteps=n)

    def macro(self):
        return self.o.macro()

    def courant(self, dt):
        return self.o.courant(dt)

    def wall(self, p):
        return self.o.wall(p)


class _EngineBackend(Engine):
    def __init__(self, case, storage="fp64"):
        super().__init__(case, storage=storage)
        self.steps = []

    def step(self, dt=None, nsteps=1):
        self.steps.append(dt)
        super().step(dt, nsteps)


def _run(path, snaps, backend_of):
    s = solver.Solver(path, backend=backend_of(foam.read_case(path)), log=None)
    out = {}
    for n in snaps:
        s.run(max_steps=n, write=False)
        out[n] = {k: v.copy() for k, v in s.fields().items()}
        out[n]["rho_w"] = s.b.wall(0)["rho_w"].copy()
    return s, out


def _compare(name, path, snaps, storage="fp64", tol=None):
    _say(f"[foam] {name} storage={storage} snaps={snaps}")
    so, ro = _run(path, snaps, _OracleBackend)
    se, re = _run(path, snaps, lambda c: _EngineBackend(c, storage))
    if tol is None:
        sr, rr = _run(path, snaps, lambda c: _OracleBackend(c, "rev"))
        sx, rx = _run(path, snaps, lambda c: _OracleBackend(c, "ext"))
    try:
        # FP32 storage moves the face velocities of the Courant number by ~1e-9
        assert np.allclose(se.b.steps, so.b.steps, rtol=1e-13 if storage == "fp64" else 1e-8, atol=0)
        for n in snaps:
            for k in ("rho", "U", "T", "q", "rho_w"):
                r = rel_l2(re[n][k], ro[n][k])
                if tol is not None:
                    _say(f"  {name} step {n} {k}: {r:.2e} (bar {tol:.1e})")
                    assert r <= tol, (name, n, k, r)
                    continue
                flx = rel_l2(rx[n][k], ro[n][k])
                floor = max(rel_l2(rr[n][k], ro[n][k]), flx)
                r_ext = rel_l2(re[n][k], rx[n][k])
                _say(f"  {name} step {n} {k}: {r:.2e} (ext {r_ext:.2e}, oracle-ext {flx:.2e}, floor {floor:.2e})")
                assert parity_ok(r, r_ext, floor, flx), (name, n, k, r, r_ext, floor, flx)
    finally:
        se.b.close()


@pytest.fixture(scope="module")
def demo_dir(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("demo_cavity"))
    return foam_box.write_fixture_case(d, os.path.join(GOLDEN, "demo_cavity_case.npz"))


def test_demo_cavity_fp64(demo_dir):
    _compare("demo_cavity", demo_dir, (1, 10, 60))


def test_demo_symmetry_fp64(tmp_path):
    """demo/testSymmetry/symmetry: half cavity with a symmetryMod plane."""
    d = foam_box.write_fixture_case(str(tmp_path), os.path.join(GOLDEN, "demo_symmetry_case.npz"))
    _compare("demo_symmetry", d, (1, 10, 40))


def test_demo_cavity_fp32_storage(demo_dir):
    _compare("demo_cavity", demo_dir, (1, 10, 60), storage="fp32", tol=1e-5)


def test_channel_case_dir_fp64(tmp_path):
    c = cases.channel(nx=24, ny=10, Lx=1.5, nv=12)
    c.rho0 = 1.0 + 0.05 * np.sin(np.linspace(0, 6, c.n_cells))
    foam_box.channel_case(str(tmp_path), c, nonuniform_rho=c.rho0)
    _compare("channel_dir", str(tmp_path), (1, 15))


def test_solver_writes_time_directories(demo_dir, tmp_path):
    """The application loop on the engine writes OpenFOAM volume fields that read back."""
    d = foam_box.write_fixture_case(str(tmp_path), os.path.join(GOLDEN, "demo_cavity_case.npz"),
                                    controls=dict(writeInterval=5))
    s = solver.Solver(d, log=None)
    try:
        s.run(max_steps=10)
        case = s.case
        sizes = {b["name"]: b["nFaces"] for b in case.meta["boundary"]}
        tn = solver.time_name(s.t, case.meta["controls"]["timePrecision"])
        T, tb = foam.read_field(os.path.join(d, tn, "T"), case.n_cells, sizes, vector=False)
        np.testing.assert_array_equal(T, s.fields()["T"])
        assert tb["movingWall"]["type"] == "fixedValue"
        assert len([x for x in os.listdir(d) if x[0].isdigit() and x != "0"]) == 2
    finally:
        s.b.close()
