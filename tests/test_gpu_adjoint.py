"""The discrete adjoint on the GPU, through the C-ABI (dugks_adjoint_*).

Parity: the GPU sweep runs the same per-point code as the host build (tests/adjoint_host.cpp), so
J and dJ/dT_w agree with it to round-off (contraction and the order of the workgroup sums:
1e-12 on J, 1e-10 on the gradient).  Its forward trajectory matches the oracle at 1e-12, and its
gradient matches central differences of the oracle's objective at 1e-6 (observed ~1e-8), along a
random direction of all wall temperatures with random weights on every face.
"""
import numpy as np
import pytest

from adjoint_cases import CASES, directional_fd, objective, random_problem, wall_mask
from adjoint_host import HostAdjoint
from dugksfoam_amd import cases, sensitivity
from dugksfoam_amd.adjoint import Adjoint
from dugksfoam_amd.engine import DugksError
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make,steps,first,every", CASES, ids=[c[0] for c in CASES])
def test_adjoint_matches_host_build_and_oracle(name, make, steps, first, every):
    print(f"adjoint case {name} steps {steps} first {first} every {every}", flush=True)
    c = make()
    coef, v = random_problem(c)
    with Adjoint(c) as A:
        J, g = A.run(steps, first, coef, checkpoint_every=every)
    Jh, gh = HostAdjoint(c).run(steps, first, coef, checkpoint_every=every)
    assert abs(J - Jh) <= 1e-12 * abs(Jh)
    assert np.linalg.norm(g - gh) <= 1e-10 * np.linalg.norm(gh)
    assert np.all(g[~wall_mask(c)] == 0.0)
    Jo = objective(Oracle, c, steps, first, coef)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    fd = directional_fd(Oracle, c, steps, first, coef, v)
    assert abs(float(g @ v) - fd) <= 1e-6 * abs(fd), (float(g @ v), fd)


def test_adjoint_forward_matches_oracle():
    c = cases.channel(nx=16, ny=6, Lx=1.0, nv=10, kn=0.5, n_modes=4)
    with Adjoint(c) as A:
        A.run(20, 20, np.zeros(c.n_faces), checkpoint_every=6)
        m = A.macro()
        fwd, rev = A.last_ms()
    assert fwd > 0 and rev > 0
    o = Oracle(c)
    o.init_equilibrium()
    o.step(nsteps=20)
    mo = o.macro()
    for k in ("rho", "U", "T", "q"):
        assert np.linalg.norm(m[k] - mo[k]) <= 1e-12 * np.linalg.norm(mo[k]), k


def test_mass_flow_gradient_matches_oracle_differences():
    kw = dict(nx=16, ny=6, Lx=1.0, nv=10, kn=0.5)
    a = np.array([0.01, -0.005, 0.003, 0.002])
    J, g = sensitivity.adjoint_gradient(a, steps=40, avg_last=10, checkpoint_every=7, **kw)
    Jf, gf = sensitivity.gradient(a, steps=40, avg_last=10, h=1e-6, backend=Oracle, **kw)
    print("adjoint", g, "oracle differences", gf, flush=True)
    assert abs(J - Jf) <= 1e-11 * abs(Jf)
    assert np.all(np.abs(g - gf) <= 1e-6 * np.abs(gf).max())
    # the engine's own objective agrees with the adjoint's forward sweep
    Je = sensitivity.mass_flow(cases.channel(a=a, n_modes=a.size, **kw), 40, 10)
    assert abs(J - Je) <= 1e-11 * abs(Je)


def test_checkpoint_spacing_and_reuse_are_exact():
    c = cases.cavity(5, 8)
    coef, _ = random_problem(c, seed=4)
    with Adjoint(c) as A:
        J1, g1 = A.run(9, 3, coef, checkpoint_every=1)
        J0, g0 = A.run(9, 3, coef, checkpoint_every=0)
        Js, gs = A.run(4, 1, coef)            # a shorter run on the same handle
    assert J1 == J0 and np.array_equal(g1, g0)
    with Adjoint(c) as B:
        Jf, gf = B.run(4, 1, coef)
    assert Js == Jf and np.array_equal(gs, gf)


def test_unsupported_and_invalid_inputs():
    with pytest.raises(DugksError) as e:
        c = cases.cavity(3, 5)
        c.patches[0]["type"] = 99
        Adjoint(c)
    assert e.value.code == -6 and "unknown patch type" in str(e.value)
    c = cases.cavity(3, 5)
    with Adjoint(c) as A:
        with pytest.raises(DugksError) as e:
            A.run(0, 1, np.zeros(c.n_faces))
        assert e.value.code == -1
        with pytest.raises(DugksError) as e:
            A.run(3, 1, np.zeros(c.n_faces), dt=-1.0)
        assert e.value.code == -1


def test_full_channel_gradient_matches_engine_differences():
    """The full micro-channel workload (200 x 40 cells, 32^2 velocities, 16 modes, 2000 steps,
    mid-plane mass flow over the last 500): dJ/da from the adjoint against central differences
    of the FP64 engine (h = 1e-6) for all 16 components.  Bar: the gradient vector within 1e-6
    relative L2, and every component that carries at least 1 % of the largest one within 1e-6
    relative.  A smaller component is resolved by the differences only to their round-off,
    about eps |J| / h absolute (J itself is a sum rounded in FP64 at every step): mode 5, about
    2000 times below mode 1, differs by 2.6e-11 absolute, which is that level."""
    c = cases.config("channel")
    a = np.asarray(c.meta["a"], dtype=np.float64)
    kw = dict(nx=c.meta["nx"], ny=c.meta["ny"], Lx=c.meta["Lx"], kn=c.meta["kn"])
    print("adjoint on the full channel", flush=True)
    J, g = sensitivity.adjoint_gradient(a, steps=2000, avg_last=500, checkpoint_every=100, **kw)
    Je = sensitivity.mass_flow(cases.channel(a=a, n_modes=a.size, **kw), 2000, 500)
    assert abs(J - Je) <= 1e-10 * abs(Je)
    print(f"J = {Je:.12e}; eps |J| / h = {np.finfo(float).eps * abs(Je) / 1e-6:.1e}", flush=True)
    fd = np.zeros_like(g)
    for k in range(a.size):
        ap, am = a.copy(), a.copy()
        ap[k] += 1e-6
        am[k] -= 1e-6
        Jp = sensitivity.mass_flow(cases.channel(a=ap, n_modes=a.size, **kw), 2000, 500)
        Jm = sensitivity.mass_flow(cases.channel(a=am, n_modes=a.size, **kw), 2000, 500)
        fd[k] = (Jp - Jm) / 2e-6
        r = abs(g[k] - fd[k]) / abs(fd[k])
        print(f"mode {k + 1}: adjoint {g[k]:.10e} engine differences {fd[k]:.10e} rel {r:.1e} "
              f"abs {abs(g[k] - fd[k]):.1e}", flush=True)
    rv = float(np.linalg.norm(g - fd) / np.linalg.norm(fd))
    print(f"gradient vector: relative L2 {rv:.2e}", flush=True)
    assert rv <= 1e-6, rv
    big = np.abs(fd) >= 1e-2 * np.abs(fd).max()
    bad = [(k + 1, g[k], fd[k]) for k in range(a.size) if big[k] and not abs(g[k] - fd[k]) <= 1e-6 * abs(fd[k])]
    assert not bad, bad

