"""Mass-flow objective of the micro-channel (dugksfoam_amd.sensitivity).

CPU (oracle backend): J vanishes for a uniform wall temperature (gas at rest, no driving) and is
odd in the coefficients (the reflection x -> L - x maps T(x; a) to T(x; -a) and reverses the
flow through the mid plane).  The torch.autograd.Function: backward() gives the adjoint
gradient (the host build of the adjoint on the CPU, the HIP adjoint on the GPU), which matches
the oracle's central differences.  GPU: the engine's central-difference gradient equals the
oracle's.
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


def _oracle_fd(oracle_lib, a, steps, avg_last):
    return sensitivity.gradient(a, steps=steps, avg_last=avg_last, h=1e-6, backend=_oracle_backend(oracle_lib),
                                **SMALL)


def test_autograd_function_on_cpu(oracle_lib):
    """mass_flow_function's backward() with the host build of the adjoint: equal to
    adjoint_gradient, and within 1e-6 of the oracle's central differences; J equals the
    oracle's mass flow; J and the gradient keep the coefficients' dtype and device."""
    import torch
    from adjoint_host import HostAdjoint
    a0 = np.array([0.02, -0.01, 0.015])
    steps, avg = 12, 4
    fn = sensitivity.mass_flow_function(steps, avg, backend=HostAdjoint, **SMALL)
    a = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
    J = fn.apply(a)
    J.backward()
    assert J.dtype == torch.float64 and J.device == a.device and a.grad.device == a.device
    Ja, ga = sensitivity.adjoint_gradient(a0, steps, avg, backend=HostAdjoint, **SMALL)
    np.testing.assert_array_equal(a.grad.numpy(), ga)
    Jo, go = _oracle_fd(oracle_lib, a0, steps, avg)
    assert abs(float(J) - Jo) <= 1e-12 * abs(Jo)
    assert np.linalg.norm(a.grad.numpy() - go) <= 1e-6 * np.linalg.norm(go)
    # a scaled output gradient scales the input gradient
    a.grad = None
    (3.0 * fn.apply(a)).backward()
    np.testing.assert_allclose(a.grad.numpy(), 3.0 * ga, rtol=1e-15)


_GPU_CHILD = r"""
import sys
import numpy as np
import torch
# torch's HIP runtime first: the engine library then binds to the same runtime (a process whose
# first HIP user is the engine library keeps torch from seeing the device)
torch.ones(1, device="cuda")
sys.path.insert(0, sys.argv[1])
from dugksfoam_amd import sensitivity
SMALL = dict(nx=16, ny=6, Lx=1.0, nv=10, kn=0.5)
a0 = np.array([0.02, -0.01, 0.015])
fn = sensitivity.mass_flow_function(12, 4, **SMALL)
a = torch.tensor(a0, dtype=torch.float64, device="cuda", requires_grad=True)
J = fn.apply(a)
J.backward()
assert J.device == a.device and a.grad.device == a.device, (J.device, a.grad.device)
Ja, ga = sensitivity.adjoint_gradient(a0, 12, 4, **SMALL)
np.testing.assert_array_equal(a.grad.cpu().numpy(), ga)
print(repr(float(J.detach())), " ".join(repr(float(x)) for x in a.grad.cpu().numpy()), flush=True)
"""


@pytest.mark.gpu
def test_autograd_function_on_gpu(oracle_lib):
    """The same with the HIP adjoint and the coefficients on the GPU: J and dJ/da come back on
    the GPU, backward() equals adjoint_gradient and the oracle's central differences (1e-6).
    Runs in a child process that initialises torch's device before the engine library loads,
    so that torch and the library share one HIP runtime."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _GPU_CHILD, root], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    vals = [float(x) for x in r.stdout.strip().splitlines()[-1].split()]
    J, g = vals[0], np.array(vals[1:])
    Jo, go = _oracle_fd(oracle_lib, np.array([0.02, -0.01, 0.015]), 12, 4)
    print("J", J, Jo, "grad", g, go, flush=True)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert np.linalg.norm(g - go) <= 1e-6 * np.linalg.norm(go)
