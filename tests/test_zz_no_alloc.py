"""dugks_get_df_noalloc (the allocation-free variant of dugks_get_df) against dugks_get_df.

Both read the same resident state through the C-ABI; the no-allocation entry point passes the
rows through the library's fixed per-GPU scratch array.  Two calls in a row, the second with a
smaller cell list, must each equal dugks_get_df bit for bit (the same kernel arithmetic); a list
beyond the scratch capacity is rejected with DUGKS_ERR_INVALID.
"""
import sys

import numpy as np
import pytest

from dugksfoam_amd import cases
from dugksfoam_amd.engine import DugksError, Engine

pytestmark = pytest.mark.gpu

MAX_DOUBLES = 1 << 20   # DUGKS_DF_NOALLOC_MAX_DOUBLES


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
@pytest.mark.parametrize("mk", [lambda: cases.cavity(10, 13), lambda: cases.cavity(5, 8, dim=3, K=2),
                                lambda: cases.cavity(6, 12, dim=3, kn=0.1)],
                         ids=["cavity2d", "cavity3d_K2", "cavity3d_noh"])
def test_df_noalloc_matches_get_df(mk, storage):
    c = mk()
    _say(f"[noalloc] cells={c.n_cells} nv={c.nv} storage={storage}")
    e = Engine(c, storage=storage)
    e.init_equilibrium()
    e.step(nsteps=3)
    rng = np.random.default_rng(3)
    big = rng.choice(c.n_cells, size=min(40, c.n_cells), replace=False).astype(np.int32)
    small = big[:5][::-1].copy()
    for cells in (big, small):
        g0, h0 = e.df(cells)
        g1, h1 = e.df_noalloc(cells)
        assert np.array_equal(g0, g1), (cells.size, float(np.abs(g0 - g1).max()))
        assert np.array_equal(h0, h1), (cells.size, float(np.abs(h0 - h1).max()))
    # the state keeps advancing: later calls still read the current state
    e.step(nsteps=1)
    g0, _ = e.df(small)
    g1, _ = e.df_noalloc(small)
    assert np.array_equal(g0, g1)
    e.close()


def test_df_noalloc_capacity():
    """Cell lists larger than the scratch array are refused; the largest one that fits works."""
    c = cases.cavity(24, 41)            # nv = 41^2 = 1681: at most 311 cells
    e = Engine(c)
    e.init_equilibrium()
    e.step(nsteps=2)
    nmax = 0
    while 2 * (nmax + 1) * c.nv + ((4 * (nmax + 1) + 255) // 256) * 32 <= MAX_DOUBLES:
        nmax += 1
    assert nmax == 311
    cells = np.arange(nmax, dtype=np.int32)
    g0, h0 = e.df(cells)
    g1, h1 = e.df_noalloc(cells)
    assert np.array_equal(g0, g1) and np.array_equal(h0, h1)
    with pytest.raises(DugksError) as x:
        e.df_noalloc(np.arange(nmax + 1, dtype=np.int32))
    assert x.value.code == -1
    # a refused call leaves the engine usable
    g2, _ = e.df_noalloc(cells[:3])
    assert np.array_equal(g2, g0[:3])
    e.close()
