"""Mass-flow objective of the micro-channel (dugksfoam_amd.sensitivity).

CPU (oracle backend): J vanishes for a uniform wall temperature (gas at rest, no driving) and is
odd in the coefficients (the reflection x -> L - x maps T(x; a) to T(x; -a) and reverses the
flow through the mid plane).  GPU: the engine's central-difference gradient equals the oracle's.
"""
import numpy as np
import pytest

from dugksfoam_amd import sensitivity

SMALL = dict(nx=16, ny=6, Lx=1.0, nv=10, kn=0.5)


def _oracle_backend(oracle_lib):
    def make(case):
        o = oracle_lib.Oracle(case, threads=4)
        return o
    return make


def test_mass_flow_zero_and_odd(oracle_lib):
    from dugksfoam_amd import cases
    be = _oracle_backend(oracle_lib)
    a = np.array([0.02, -0.01, 0.015])
    J0 = sensitivity.mass_flow(cases.channel(a=np.zeros(3), n_modes=3, **SMALL), 12, 4, be)
    Jp = sensitivity.mass_flow(cases.channel(a=a, n_modes=3, **SMALL), 12, 4, be)
    Jm = sensitivity.mass_flow(cases.channel(a=-a, n_modes=3, **SMALL), 12, 4, be)
    assert abs(Jp) > 0
    assert abs(J0) <= 1e-12 * abs(Jp)
    assert abs(Jp + Jm) <= 1e-9 * abs(Jp)


@pytest.mark.gpu
def test_engine_gradient_matches_oracle(oracle_lib):
    a = np.array([0.02, -0.01, 0.015, 0.005])
    kw = dict(steps=40, avg_last=10, h=1e-6, **SMALL)
    Je, ge = sensitivity.gradient(a, **kw)
    Jo, go = sensitivity.gradient(a, backend=_oracle_backend(oracle_lib), **kw)
    print("J", Je, Jo, "grad", ge, go, flush=True)
    assert abs(Je - Jo) <= 1e-11 * abs(Jo)
    assert np.linalg.norm(ge - go) <= 1e-5 * np.linalg.norm(go)
