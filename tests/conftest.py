import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the HIP engine)")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


FP64_TOL = 1e-12       # FP64 parity bar, relative L2
FLOOR_FACTOR = 2.0


def parity_ok(err_oracle, err_ext, floor, floor_ext=0.0):
    """The FP64 parity bar.  Either the engine is within 1e-12 of the FP64 oracle, or the field
    is one whose sums cancel almost to zero (the velocity one step after rest, a slow
    transpiration flow, the heat flux of cells along two walls), where the reference's FP64
    arithmetic is itself that far from the exact result.  Then two checks must hold:
      - the engine is within 1e-12 of the extended-precision oracle (80-bit intermediates and
        velocity sums; err_ext, None when it was not evaluated), or, where the FP64 oracle is
        itself farther than that from it (floor_ext: their distance), no farther from it than
        the FP64 oracle is: no FP64 evaluation of the step can do better;
      - the engine is within FLOOR_FACTOR floors of the FP64 oracle (floor: the FP64 oracle's
        distance from the extended-precision and reversed-sum oracles).
    A difference of the step itself fails both."""
    if err_oracle <= FP64_TOL:
        return True
    if err_ext is None:
        return False
    return err_ext <= max(FP64_TOL, floor_ext) and err_oracle <= FLOOR_FACTOR * floor


def seq_mass(V, rho):
    """Total mass summed cell by cell in cell order, as the oracle's total_mass sums it."""
    return float(np.cumsum(np.asarray(V) * np.asarray(rho))[-1])


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as orc
    orc.build()
    return orc
