"""The discrete adjoint's own code (adjoint_core.h, adjoint_driver.h), run on the host, against the
oracle: its forward trajectory against the oracle's steps, its gradient against central
differences of the oracle's objective.

The adjoint is pinned only by the oracle: dJ/dT_w along a random direction of all wall
temperatures must equal the central difference of the oracle's J (h = 1e-3 K) with random
weights on every face, so every face kind, branch and the wall density path enter J.  The
tolerance 1e-6 is the one the adjoint is specified to; the observed agreement is about 1e-8.
"""
import numpy as np
import pytest

from adjoint_cases import CASES, directional_fd, objective, random_problem, wall_mask
from adjoint_host import HostAdjoint
from dugksfoam_amd import cases, sensitivity
from oracle.oracle import Oracle


@pytest.mark.parametrize("name,make,steps,first,every", CASES, ids=[c[0] for c in CASES])
def test_directional_derivative(name, make, steps, first, every):
    c = make()
    coef, v = random_problem(c)
    J, g = HostAdjoint(c).run(steps, first, coef, checkpoint_every=every)
    Jo = objective(Oracle, c, steps, first, coef)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    fd = directional_fd(Oracle, c, steps, first, coef, v)
    ad = float(g @ v)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    assert np.all(g[~wall_mask(c)] == 0.0)


def test_forward_matches_oracle():
    c = cases.channel(nx=16, ny=6, Lx=1.0, nv=10, kn=0.5, n_modes=4)
    H = HostAdjoint(c)
    H.run(12, 12, np.zeros(c.n_faces))
    o = Oracle(c)
    o.init_equilibrium()
    o.step(nsteps=12)
    mo, mh = o.macro(), H.macro()
    for k in ("rho", "U", "T", "q"):
        assert np.linalg.norm(mh[k] - mo[k]) <= 1e-12 * np.linalg.norm(mo[k]), k


def test_mass_flow_gradient_matches_oracle_differences():
    """dJ/da of the channel's mass-flow objective, chained from dJ/dT_w, against central
    differences of the oracle in the Fourier coefficients."""
    kw = dict(nx=16, ny=6, Lx=1.0, nv=10, kn=0.5)
    a = np.array([0.01, -0.005, 0.003, 0.002])
    J, g = sensitivity.adjoint_gradient(a, steps=12, avg_last=4, backend=HostAdjoint, **kw)
    Jf, gf = sensitivity.gradient(a, steps=12, avg_last=4, h=1e-6, backend=Oracle, **kw)
    assert abs(J - Jf) <= 1e-12 * abs(Jf)
    assert np.all(np.abs(g - gf) <= 1e-6 * np.abs(gf).max()), (g, gf)


def test_checkpoint_spacing_does_not_change_results():
    c = cases.cavity(4, 6)
    coef, _ = random_problem(c, seed=3)
    H = HostAdjoint(c)
    ref = H.run(7, 2, coef, checkpoint_every=1)
    for every in (0, 2, 3, 7, 50):
        J, g = H.run(7, 2, coef, checkpoint_every=every)
        assert J == ref[0] and np.array_equal(g, ref[1]), every


def test_planted_mistake_is_caught():
    """A wrong sign in the objective's weight changes J, and a perturbed wall temperature moves
    the difference quotient away from the adjoint: the comparison has the power to fail."""
    c = cases.cavity(4, 6)
    coef, v = random_problem(c, seed=5)
    J, g = HostAdjoint(c).run(4, 1, coef)
    fd_wrong = directional_fd(Oracle, c, 4, 1, -coef, v)
    assert abs(float(g @ v) - fd_wrong) > 1e-3 * abs(fd_wrong)


def test_unknown_patch_type_rejected():
    c = cases.cavity(3, 5)
    c.patches[0]["type"] = 99
    with pytest.raises(RuntimeError, match="unknown patch type"):
        HostAdjoint(c)
