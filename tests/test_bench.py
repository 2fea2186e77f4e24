"""bench.py's contract pieces that need no GPU: the refusal of several GPUs, the algorithmic
bytes behind roofline.achieved (48 B per cell-velocity update in FP64, 24 B in FP32), the
engine's own design byte model, the committed rocprof traffic and the host description."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402
from dugksfoam_amd import cases  # noqa: E402


def test_several_gpus_are_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_byte_model_of_the_3d_cavity():
    """FP64, h elided (D = 3, K = 0), face values stored: the scheme's floor is gbar read twice
    and written once plus the face values written and read once, 76 B per cell-velocity update;
    with the fused cell tail k_cell_fin and k_cell_s1 are not separate kernels."""
    c = cases.config("cavity3d")
    nc, nv = c.n_cells, c.nv
    kb, step = bench.bytes_model(c, 8, 1, 1, face_values=True, fused_cell=True)
    assert set(kb) == {"k_face_values", "k_face_moments", "k_face_fin", "k_cell"}
    assert round(step / (nc * nv), 1) == 76.0
    faces = c.n_faces
    assert kb["k_face_moments"] >= 8 * faces * nv            # every face value read once
    assert kb["k_cell"] >= 8 * (2 * nc * nv + faces * nv)     # gbar read and written, face values read
    kb2, step2 = bench.bytes_model(c, 8, 1, 1, face_values=True, fused_cell=False)
    assert {"k_cell_fin", "k_cell_s1"} <= set(kb2) and step2 == step
    assert kb2["k_cell_s1"] == 8 * 2 * nc * nv
    kb4, step4 = bench.bytes_model(c, 4, 1, 1, face_values=True, fused_cell=True)
    assert step4 * 2 == pytest.approx(step, rel=1e-3)        # FP32 storage halves the stream


def test_committed_traffic_is_above_the_algorithmic_bytes():
    c = cases.config("cavity3d")
    kb, _ = bench.bytes_model(c, 8, 1, 1, face_values=True, fused_cell=True)
    t = bench.traffic_from_profile("k_cell", "fp64")
    assert t is not None and t >= kb["k_cell"]


@pytest.mark.parametrize("wl", ["cavity2d", "delaunay", "batch", "channel", "cavity3d_share"])
def test_every_workload_has_step_traffic_above_the_contract(wl):
    """Each workload's committed counter file covers the four step kernels in both storage forms,
    and the measured bytes of a step are at least the contract's 6 s bytes per update."""
    c = cases.config(wl)
    kernels = ["k_face_values", "k_face_moments", "k_face_fin", "k_cell"]
    for st, s in (("fp64", 8), ("fp32", 4)):
        t = bench.step_traffic_from_profile(st, kernels, wl)
        assert t is not None, (wl, st)
        assert t >= bench.contract_bytes_per_update(s) * c.n_cells * c.nv, (wl, st)


def test_contract_bytes_per_update():
    """roofline.achieved counts B = 6 s per update: g and h streamed in twice and out once."""
    assert bench.contract_bytes_per_update(8) == 48
    assert bench.contract_bytes_per_update(4) == 24


def test_step_traffic_sums_the_step_kernels():
    kernels = ["k_face_values", "k_face_moments", "k_face_fin", "k_cell"]
    t = bench.step_traffic_from_profile("fp64", kernels)
    assert t == sum(bench.traffic_from_profile(k, "fp64") for k in kernels)
    assert bench.step_traffic_from_profile("fp64", kernels + ["no_such_kernel"]) is None


def test_host_description():
    h = bench.host_description(3)
    assert h["threads_used"] == 3 and h["nproc"] >= 1
    assert "cpu_model" in h and h["ram_GiB"] > 0
