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


def parity_ok(err_oracle, err_ext, floor):
    """The FP64 parity bar.  Either the engine is within 1e-12 of the FP64 oracle, or the field
    is one whose sums cancel almost to zero (the velocity one step after rest, a slow
    transpiration flow, the heat flux of cells along two walls), where the reference's
    sequential FP64 sums are themselves that far from the exact sums: then the engine must be
    within 1e-12 of the extended-precision oracle (80-bit intermediates and velocity sums;
    err_ext, None when it was not evaluated) and within FLOOR_FACTOR floors of the FP64 one
    (floor: the FP64 oracle's distance from the extended-precision and reversed-sum oracles).
    A difference of the step itself fails both."""
    if err_oracle <= FP64_TOL:
        return True
    return err_ext is not None and err_ext <= FP64_TOL and err_oracle <= FLOOR_FACTOR * floor


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as orc
    orc.build()
    return orc
