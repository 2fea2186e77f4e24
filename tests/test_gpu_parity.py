"""GPU parity: the HIP engine (through the C-ABI) against the FP64 oracle.

FP64 bar: relative L2 <= 1e-12 on rho, U, T, q (cells and faces) and the wall
data; the total-mass history equal to the oracle's to round-off.  FP32-storage
bar: the tolerances stated per test (storage rounding of g/h ~ 6e-8).
"""
import os
import sys

import numpy as np
import pytest

from conftest import FP64_TOL, GOLDEN, parity_ok, rel_l2, seq_mass

FP32_TOL = 1e-5   # FP32 storage against the FP64 oracle, relative L2 per field
from dugksfoam_amd import cases
from dugksfoam_amd.engine import DugksError, Engine

pytestmark = pytest.mark.gpu



def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle(c):
    from oracle.oracle import Oracle
    o = Oracle(c)
    o.init_equilibrium()
    return o


def _small_cases():
    return {
        "cavity2d_12x21": (lambda: cases.cavity(12, 21), 12),
        "cavity3d_6x12": (lambda: cases.cavity(6, 12, dim=3, kn=0.1), 4),
        "cavity3d_K2": (lambda: cases.cavity(5, 8, dim=3, K=2), 3),
        "channel_20x8": (lambda: cases.channel(20, 8, 1.25, 0.5, nv=12), 10),
        "delaunay_300": (lambda: cases.delaunay_cavity(300, 12, nv=12), 6),
        "batch_4": (lambda: cases.batch(n_cases=4, nmin=4, nmax=8, nv=8, seed=5), 5),
        "symcavity_6x12": (lambda: cases.symmetry_cavity(6, 12, nv=12), 10),
        "symcavity_odd": (lambda: cases.symmetry_cavity(5, 10, nv=9), 8),
        "symcavity_left": (lambda: cases.symmetry_cavity(5, 10, nv=10, plane="x0"), 8),
        "symplane_6x12": (lambda: cases.symmetry_cavity(6, 12, nv=12, kind="symmetryPlane"), 10),
        "symplane_odd": (lambda: cases.symmetry_cavity(5, 10, nv=9, kind="symmetryPlane"), 8),
        "symplane_left": (lambda: cases.symmetry_cavity(5, 10, nv=10, plane="x0", kind="symmetryPlane"), 8),
        "openchannel_12x6": (lambda: cases.open_channel(12, 6, nv=12), 10),
        "pchannel_12x6": (lambda: cases.pressure_channel(12, 6, nv=12), 10),
        "zgchannel_12x6": (lambda: cases.open_channel(12, 6, nv=12, outlet="zeroGradient"), 10),
        "agglom_b3": (lambda: cases.agglomerated_cavity(12, nv=12, block=3), 8),
        "agglom_b4": (lambda: cases.agglomerated_cavity(12, nv=9, block=4, K=1), 8),
        "fourier_1d": (lambda: cases.fourier_flow(12, nv=13), 10),
        "shocktube_1d": (lambda: cases.shock_tube(16, nv=12), 8),
        "shocktube_1d_wall": (lambda: cases.shock_tube(12, nv=10, ends="wall", K=1), 8),
    }


@pytest.fixture(scope="module", params=list(_small_cases()))
def pair(request):
    mk, n = _small_cases()[request.param]
    c = mk()
    _say(f"[parity] case {request.param} cells={c.n_cells} nv={c.nv} steps={n}")
    o = _oracle(c)
    e = Engine(c)
    e.init_equilibrium()
    mass_o, mass_e = [], []
    for _ in range(n):
        o.step()
        e.step()
        mass_o.append(o.total_mass())
        mass_e.append(seq_mass(c.V, e.macro()["rho"]))
    yield request.param, c, o, e, np.array(mass_o), np.array(mass_e)
    e.close()
    o.close()


def test_cell_fields_fp64(pair):
    name, c, o, e, _, _ = pair
    mo, me = o.macro(), e.macro()
    for k in ("rho", "U", "T", "q"):
        r = rel_l2(me[k], mo[k])
        _say(f"  {name} cell {k}: {r:.2e}")
        assert r <= FP64_TOL, (name, k, r)


def test_face_fields_fp64(pair):
    name, c, o, e, _, _ = pair
    fo, fe = o.face_macro(), e.face_macro()
    for k in ("rho", "U", "T", "q"):
        r = rel_l2(fe[k], fo[k])
        assert r <= FP64_TOL, (name, k, r)


def test_wall_data_fp64(pair):
    """rho_w at 1e-12; qWall and stressWall (fvDVM.C:539-581, diagnostics outside the
    compared fields) are velocity sums of terms ~ rho c^3 that cancel to a small
    wall heat flux, so their rounding is compared at 1e-10."""
    name, c, o, e, _, _ = pair
    tol = dict(rho_w=FP64_TOL, q_wall=1e-10, stress_wall=1e-10)
    for i, p in enumerate(c.patches):
        if p["type"] != cases.PATCH_WALL:
            continue
        wo, we = o.wall(i), e.wall(i)
        for k in ("rho_w", "q_wall", "stress_wall"):
            r = rel_l2(we[k], wo[k])
            assert r <= tol[k], (name, p["name"], k, r)


def test_mass_history(pair):
    name, c, o, e, mo, me = pair
    np.testing.assert_allclose(me, mo, rtol=1e-13)


def test_distribution_functions(pair):
    name, c, o, e, _, _ = pair
    cells = np.unique(np.linspace(0, c.n_cells - 1, 7).astype(np.int32))
    go, ho = o.df(cells)
    ge, he = e.df(cells)
    assert np.abs(ge - go).max() <= 1e-12 * np.abs(go).max()
    if np.abs(ho).max() > 0:
        assert np.abs(he - ho).max() <= 1e-12 * np.abs(ho).max()
    else:
        assert np.abs(he).max() == 0.0


def test_courant(pair):
    name, c, o, e, _, _ = pair
    a, b = o.courant(), e.courant()
    assert b[0] == pytest.approx(a[0], rel=1e-12)
    assert b[1] == pytest.approx(a[1], rel=1e-12)


@pytest.mark.parametrize("which", ["cavity2d_12x21", "cavity3d_6x12", "delaunay_300", "channel_20x8", "symcavity_6x12",
                                   "symplane_6x12", "openchannel_12x6", "zgchannel_12x6", "fourier_1d", "shocktube_1d",
                                   "agglom_b3"])
def test_fp32_storage(which):
    """FP32 storage of g/h (FP64 arithmetic) against the FP64 oracle: every field within 1e-5."""
    mk, n = _small_cases()[which]
    c = mk()
    _say(f"[fp32] case {which}")
    o = _oracle(c)
    o.step(nsteps=n)
    e = Engine(c, storage="fp32")
    e.init_equilibrium()
    e.step(nsteps=n)
    mo, me = o.macro(), e.macro()
    errs = {k: rel_l2(me[k], mo[k]) for k in mo}
    _say("  ", errs)
    for k, v in errs.items():
        assert v <= FP32_TOL, (which, k, v)
    e.close()


BATCH_SEL = list(range(0, 256, 16))   # 16 of the 256 cases of the batch workload (seed 1)


def _per_case(c, m, ref):
    """Relative L2 of each field over each case's cells: {case: {field: err}}."""
    out = {}
    for i in range(c.n_cases):
        sel = c.cell_case == i
        out[i] = {k: rel_l2(m[k][sel], ref[k][sel]) for k in ("rho", "U", "T", "q")}
    return out


def test_batch_fp32_against_oracle():
    """The batch workload's form: FP32 storage, per-case dt and muRef, here 16 of its 256 cases
    (sides 23..79, Kn 0.011..2.2), against the FP64 oracle run on the same union for 40 steps.
    Every case's rho, U, T and q within 1e-5."""
    c = cases.batch(indices=BATCH_SEL)
    _say(f"[batch-fp32] cases={c.n_cases} cells={c.n_cells} nv={c.nv}")
    from oracle.oracle import Oracle
    o = Oracle(c, threads=16)
    o.init_equilibrium()
    o.step(nsteps=40)
    e = Engine(c, storage="fp32")
    e.init_equilibrium()
    e.step(nsteps=40)
    errs = _per_case(c, e.macro(), o.macro())
    worst = {k: max(errs[i][k] for i in errs) for k in ("rho", "U", "T", "q")}
    _say("  worst case per field after 40 steps: " + " ".join(f"{k}={v:.2e}" for k, v in worst.items()))
    e.close()
    bad = [(i, k, v) for i in errs for k, v in errs[i].items() if not v <= FP32_TOL]
    assert not bad, bad


def test_batch_fp32_full_steps():
    """The same 16 cases for the workload's 1000 steps: FP32 storage against the FP64 engine
    (itself at the oracle's parity bar, test_golden_full_size[batch] and the small batch), every
    case's rho, U, T and q within 1e-5."""
    c = cases.batch(indices=BATCH_SEL)
    _say(f"[batch-fp32-1000] cases={c.n_cases} cells={c.n_cells}")
    ref = Engine(c)
    ref.init_equilibrium()
    ref.step(nsteps=1000)
    e = Engine(c, storage="fp32")
    e.init_equilibrium()
    e.step(nsteps=1000)
    errs = _per_case(c, e.macro(), ref.macro())
    worst = {k: max(errs[i][k] for i in errs) for k in ("rho", "U", "T", "q")}
    _say("  worst case per field after 1000 steps: " + " ".join(f"{k}={v:.2e}" for k, v in worst.items()))
    e.close()
    ref.close()
    bad = [(i, k, v) for i in errs for k, v in errs[i].items() if not v <= FP32_TOL]
    assert not bad, bad


def test_dt_change_and_steps_split():
    """dt changes between calls (adjustTimeStep): the resident state is re-prepared."""
    c = cases.cavity(8, 13)
    o = _oracle(c)
    e = Engine(c)
    e.init_equilibrium()
    for dt, n in ((c.dt, 3), (0.7 * c.dt, 2), (c.dt, 1)):
        o.step(dt, n)
        e.step(dt, n)
    mo, me = o.macro(), e.macro()
    for k in mo:
        assert rel_l2(me[k], mo[k]) <= FP64_TOL, k


@pytest.mark.parametrize("mk", [lambda: cases.cavity(10, 13), lambda: cases.pressure_channel(10, 5, nv=10)],
                         ids=["cavity", "pressure_channel"])
def test_save_load_resume_bit_exact(mk):
    """The state image holds everything a step reads, including the boundary state that
    pressure patches carry from step to step."""
    c = mk()
    e = Engine(c)
    e.init_equilibrium()
    e.step(nsteps=4)
    img = e.save_state()
    e.step(nsteps=3)
    a = e.macro()
    e.load_state(img)
    e.step(nsteps=3)
    b = e.macro()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # a fresh engine resumes from the image too
    e2 = Engine(c)
    e2.init_equilibrium()
    e2.load_state(img)
    e2.step(nsteps=3)
    d = e2.macro()
    for k in a:
        assert np.array_equal(a[k], d[k]), k


def test_error_paths():
    c = cases.cavity(4, 5)
    e = Engine(c)
    with pytest.raises(DugksError) as x:
        e.step(nsteps=1)
    assert x.value.code == -4
    bad = c.rho0.copy()
    bad[3] = -1.0
    with pytest.raises(DugksError) as x:
        e.init_equilibrium(rho=bad)
    assert x.value.code == -1
    T = c.T0.copy()
    T[0] = np.nan
    with pytest.raises(DugksError):
        e.init_equilibrium(T=T)
    e.init_equilibrium()
    with pytest.raises(DugksError):
        e.step(dt=-1.0)
    with pytest.raises(DugksError):
        e.wall(5)
    with pytest.raises(DugksError):
        e.df([c.n_cells])
    with pytest.raises(DugksError):
        e.load_state(b"\0" * 64)
    e.step(nsteps=1)


def test_nonfinite_is_reported():
    """A time step far beyond stability drives rho/T non-finite; the step reports it."""
    c = cases.cavity(4, 9)
    e = Engine(c)
    e.init_equilibrium()
    with pytest.raises(DugksError) as x:
        e.step(dt=c.dt * 1e4, nsteps=40)
    assert x.value.code == -5


def _golden(name):
    p = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(p):
        pytest.skip(f"fixture {name}.npz not generated")
    return np.load(p)


def _golden_floor(g, k, s):
    """The FP64 oracle's own round-off for field k at step s: its larger distance from two other
    evaluations of the same arithmetic, one with 80-bit intermediates and 80-bit velocity sums
    (near-exact sums, the ext oracle), one with every velocity sum in reverse order."""
    fl = [float(g[key]) for key in (f"floor_ext_{k}_{s}", f"floor_rev_{k}_{s}") if key in g]
    return max(fl) if fl else 0.0


def _check_snap(g, s, m, failures):
    errs = {k: rel_l2(m[k], g[f"{k}_{s}"]) for k in ("rho", "U", "T", "q")}
    exts = {k: (rel_l2(m[k], g[f"ext_{k}_{s}"]) if f"ext_{k}_{s}" in g else None) for k in errs}
    fls = {k: _golden_floor(g, k, s) for k in errs}
    flx = {k: float(g[f"floor_ext_{k}_{s}"]) if f"floor_ext_{k}_{s}" in g else 0.0 for k in errs}

    def note(k, v):
        if v <= FP64_TOL:
            return ""
        ext = "not stored" if exts[k] is None else f"{exts[k]:.2e}"
        return f" (ext {ext}, oracle-ext {flx[k]:.1e}, floor {fls[k]:.1e})"

    _say(f"  step {s}: " + " ".join(f"{k}={v:.2e}" + note(k, v) for k, v in errs.items()))
    failures += [(int(s), k, v, exts[k]) for k, v in errs.items() if not parity_ok(v, exts[k], fls[k], flx[k])]


def _check_snap_sampled(g, s, m, failures):
    st = int(g["sample"]) if "sample" in g else 1
    _check_snap(g, s, {k: v[::st] for k, v in m.items()}, failures)


@pytest.mark.parametrize("name", ["cavity2d", "channel", "batch", "delaunay"])
def test_golden_full_size(name):
    """Full configurations against committed oracle fixtures: the 2-D cavity and the channel for
    their 2000 steps, the 256-case batch (FP64, per-case dt and muRef) for 2 steps, the ~200k
    triangle Delaunay cavity for 5 steps (fields of every 7th cell)."""
    g = _golden(name)
    c = cases.config(name)
    _say(f"[golden] {name} cells={c.n_cells} nv={c.nv} snaps={list(g['snaps'])}")
    e = Engine(c)
    e.init_equilibrium()
    n = 0
    mass = [seq_mass(c.V, c.rho0)]
    failures = []
    for s in g["snaps"]:
        while n < s:
            e.step(nsteps=1)
            n += 1
            mass.append(seq_mass(c.V, e.macro()["rho"]))
        _check_snap_sampled(g, s, e.macro(), failures)
    e.close()
    assert not failures, failures
    np.testing.assert_allclose(np.array(mass), g["mass"], rtol=1e-12)


_SMALLEST = {
    # one cell whose faces are all walls: no internal face, so no face on the matrix-core
    # moment kernel, and a least-squares stencil of wall rows only
    "cavity2d_1cell": lambda: cases.cavity(1, 9),
    "cavity3d_1cell": lambda: cases.cavity(1, 6, dim=3),
    "cavity2d_2x2": lambda: cases.cavity(2, 7),
    "fourier_1d_1cell": lambda: cases.fourier_flow(1, nv=9),
    "shocktube_1d_2cells": lambda: cases.shock_tube(2, nv=8),
}


@pytest.mark.parametrize("which", list(_SMALLEST))
def test_smallest_meshes(which):
    """Degenerate meshes (one or two cells, every or all but one face on a boundary) in both
    storage forms against the FP64 oracle over 5 steps: FP64 cell and face fields and the mass
    history at the FP64 bar, FP32 storage at 1e-5.  A U or q field that is itself round-off (RMS
    below 1e-10 of its natural scale, sqrt(R T) or rho (R T)^1.5) is measured against that scale
    instead: on the two walls of a one-cell Fourier flow the face velocity is the residue of equal
    in- and outgoing mass fluxes, about 1e-14 sqrt(R T) in both implementations."""
    c = _SMALLEST[which]()
    _say(f"[smallest] {which} cells={c.n_cells} faces={c.n_faces} nv={c.nv}")
    o = _oracle(c)
    e64 = Engine(c)
    e32 = Engine(c, storage="fp32")
    try:
        for e in (e64, e32):
            e.init_equilibrium()
        mo, me = [], []
        for _ in range(5):
            o.step()
            e64.step()
            e32.step()
            mo.append(o.total_mass())
            me.append(seq_mass(c.V, e64.macro()["rho"]))
        np.testing.assert_allclose(me, mo, rtol=1e-13)
        co, fo = o.macro(), o.face_macro()
        c64, f64 = e64.macro(), e64.face_macro()
        c32 = e32.macro()
        RT = c.R * float(np.mean(co["T"]))
        scale = dict(rho=0.0, T=0.0, U=np.sqrt(RT), q=float(np.mean(co["rho"])) * RT ** 1.5)

        def err(a, b, k):
            a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
            ref = scale[k] * np.sqrt(b.size)
            nb = np.linalg.norm(b)
            return float(np.linalg.norm(a - b) / (ref if nb < 1e-10 * ref else nb))

        for k in ("rho", "U", "T", "q"):
            r_c, r_f, r_32 = err(c64[k], co[k], k), err(f64[k], fo[k], k), err(c32[k], co[k], k)
            _say(f"  {which} {k}: cell {r_c:.2e} face {r_f:.2e} fp32 {r_32:.2e}")
            assert r_c <= FP64_TOL and r_f <= FP64_TOL, (which, k, r_c, r_f)
            assert r_32 <= FP32_TOL, (which, k, r_32)
    finally:
        e64.close()
        e32.close()
        o.close()


def test_delaunay_midsize():
    """An unstructured Delaunay cavity of about 10^4 triangles (irregular owner/neighbour
    gathers, non-orthogonal boundary faces, up to 8 faces per cell after smoothing) against
    the oracle over 40 steps, inside the stable window of the scheme on such meshes.  Bar:
    parity_ok, with the extended-precision and reversed-sum oracles run live (the heat flux one
    step after rest cancels)."""
    from oracle.oracle import Oracle
    c = cases.delaunay_cavity(4900, 70, nv=16)
    _say(f"[delaunay] cells={c.n_cells} nv={c.nv}")
    o = _oracle(c)
    orv = Oracle(c, threads=16, variant="rev")
    orv.init_equilibrium()
    ox = Oracle(c, threads=16, variant="ext")
    ox.init_equilibrium()
    e = Engine(c)
    e.init_equilibrium()
    for n in (1, 39):
        for x in (o, orv, ox, e):
            x.step(nsteps=n)
        mo, mr, mx, me = o.macro(), orv.macro(), ox.macro(), e.macro()
        for k in ("rho", "U", "T", "q"):
            r, rx = rel_l2(me[k], mo[k]), rel_l2(me[k], mx[k])
            flx = rel_l2(mx[k], mo[k])
            fl = max(rel_l2(mr[k], mo[k]), flx)
            _say(f"  delaunay step {n}: {k} {r:.2e} (ext {rx:.2e}, oracle-ext {flx:.2e}, floor {fl:.2e})")
            assert parity_ok(r, rx, fl, flx), (n, k, r, rx, fl, flx)
    e.close()


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_golden_cavity3d(storage):
    """48^3 x 24^3 against the committed oracle fixture (steps 1, 2, 5, 10, 20; fields of every
    7th cell): FP64 at the parity bar, FP32 storage at 1e-5 relative L2 per field."""
    g = _golden("cavity3d")
    c = cases.config("cavity3d")
    _say(f"[golden] cavity3d {storage} cells={c.n_cells} nv={c.nv} snaps={list(g['snaps'])}")
    e = Engine(c, storage=storage)
    e.init_equilibrium()
    n = 0
    failures = []
    for s in g["snaps"]:
        e.step(nsteps=int(s) - n)
        n = int(s)
        st = int(g["sample"]) if "sample" in g else 1
        m = {k: v[::st] for k, v in e.macro().items()}
        if storage == "fp64":
            _check_snap(g, s, m, failures)
        else:
            errs = {k: rel_l2(m[k], g[f"{k}_{s}"]) for k in ("rho", "U", "T", "q")}
            _say(f"  step {s}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
            failures += [(int(s), k, v) for k, v in errs.items() if not v <= FP32_TOL]
    e.close()
    assert not failures, failures


def test_delaunay_full_size_runs_its_steps():
    """The c2 workload (about 2e5 triangles x 64^2 velocities, wall row orthogonal) over its
    1000 steps, in FP64 and in FP32 storage (each field within 1e-5 of FP64): every field finite, the temperature within 250-300 K of the 273 K walls and the
    lid's flow.  The reference scheme does not conserve mass exactly (the discrete equilibrium is
    not moment-exact on the truncated velocity grid, dv.C:393-406, 937; the wall density
    balances the outgoing flux before relaxation, dv.C:623-624, while the flux uses the relaxed
    values, dv.C:905-920): the total mass drifts by about 1.2e-6 of itself over the 1000 steps,
    checked against a 1e-5 bound."""
    c = cases.config("delaunay")
    _say(f"[c2] cells={c.n_cells} nv={c.nv} steps=1000")
    e = Engine(c)
    e.init_equilibrium()
    m0 = seq_mass(c.V, c.rho0)
    for n in range(10):
        e.step(nsteps=100)
        m = e.macro()
        T = m["T"]
        _say(f"  step {100 * (n + 1)}: T {T.min():.2f}..{T.max():.2f} |U|max {np.abs(m['U']).max():.2f}")
        assert all(np.all(np.isfinite(v)) for v in m.values())
        assert 250.0 < T.min() and T.max() < 300.0
    drift = abs(seq_mass(c.V, m["rho"]) - m0) / m0
    _say(f"  mass drift {drift:.2e}")
    assert drift <= 1e-5
    e.close()
    # the FP32-storage form over the same 1000 steps, within 1e-5 of the FP64 fields
    e32 = Engine(c, storage="fp32")
    e32.init_equilibrium()
    e32.step(nsteps=1000)
    m32 = e32.macro()
    e32.close()
    for k in ("rho", "U", "T", "q"):
        r = rel_l2(m32[k], m[k])
        _say(f"  fp32 vs fp64 {k}: {r:.2e}")
        assert r <= FP32_TOL, (k, r)


def test_cavity3d_full_size_invariants():
    """48^3 x 24^3: the lid moves along x, so the flow is mirror-symmetric in z;
    the pre-relaxation wall mass flux vanishes at every wall face."""
    c = cases.config("cavity3d")
    _say(f"[invariants] cavity3d cells={c.n_cells} nv={c.nv}")
    e = Engine(c)
    e.init_equilibrium()
    e.step(nsteps=3)
    m = e.macro()
    n = 48
    r = m["rho"].reshape(n, n, n)          # [z][y][x]
    uz = m["U"][:, 2].reshape(n, n, n)
    ux = m["U"][:, 0].reshape(n, n, n)
    assert np.abs(ux).max() > 1e-3
    assert np.abs(r - r[::-1]).max() <= 1e-13 * np.abs(r).max()
    assert np.abs(uz + uz[::-1]).max() <= 1e-11 * np.abs(m["U"]).max() + 1e-12
    f = e.face_macro()
    nfi = c.n_internal_faces
    nrm = c.Sf[nfi:] / np.linalg.norm(c.Sf[nfi:], axis=1)[:, None]
    un = np.einsum("ij,ij->i", f["U"][nfi:], nrm)
    assert np.abs(un).max() <= 1e-12 * max(1.0, np.abs(f["U"]).max())
    e.close()


def test_step_graphs_match_eager():
    """Captured two-step graphs (the default for multi-step calls) give bit-identical results
    to eager launches, over several replays with host work and synchronisation between them."""
    c = cases.cavity(10, 13)
    a, b = Engine(c), Engine(c)
    a.init_equilibrium()
    b.init_equilibrium()
    b.set_graphs(False)
    for n in (4, 3, 6, 1, 2):
        a.step(nsteps=n)
        b.step(nsteps=n)
        ma, mb = a.macro(), b.macro()
        for k in ma:
            assert np.array_equal(ma[k], mb[k]), (n, k)
    a.close()
    b.close()


@pytest.mark.parametrize("which", list(_small_cases()))
def test_face_value_mode(which):
    """Face-value mode (stored face values, the default where they fit) against the mode that
    rebuilds them in the cell kernel: FP64 bit-identical when both use the same cell-kernel
    launch shape, else within the FP64 bar (the velocity sums are split differently);
    FP32 storage within the rounding of the stored face values; both directions of the
    switch on a running engine."""
    mk, n = _small_cases()[which]
    c = mk()
    if any(p["type"] not in (cases.PATCH_WALL, cases.PATCH_CYCLIC) for p in c.patches):
        # symmetry, far-field and fixed patches are handled in face-value mode only
        e = Engine(c)
        assert e.face_values()
        with pytest.raises(DugksError) as x:
            e.set_face_values(False)
        assert x.value.code == -6
        e.close()
        return
    for st, tol in (("fp64", 0.0), ("fp32", 1e-6)):
        a, b = Engine(c, storage=st), Engine(c, storage=st)
        a.set_face_values(True)
        b.set_face_values(False)
        assert a.face_values() and not b.face_values()
        if st == "fp64" and (a.cell_blocks() != b.cell_blocks() or a.face_mfma()):
            # other velocity-sum splits, or the face moments as matrix products
            tol = FP64_TOL
        a.init_equilibrium()
        b.init_equilibrium()
        a.step(nsteps=n)
        b.step(nsteps=n)
        ma, mb = a.macro(), b.macro()
        for k in ma:
            r = rel_l2(ma[k], mb[k])
            _say(f"  {which} {st} face values on/off {k}: {r:.2e}")
            if tol == 0.0:
                assert np.array_equal(ma[k], mb[k]), (which, st, k, r)
            else:
                assert r <= tol, (which, st, k, r)
        if st == "fp64":
            a.set_face_values(False)
            b.set_face_values(True)
            a.step(nsteps=3)
            b.step(nsteps=3)
            ma, mb = a.macro(), b.macro()
            for k in ma:
                assert rel_l2(ma[k], mb[k]) <= FP64_TOL, (which, "switched", k)
        a.close()
        b.close()


_BIT_CASES = {
    # velocity counts that pad to 256, so that both modes pick the same cell-kernel shape
    "cavity2d_12x16": (lambda: cases.cavity(12, 16), 6),
    "cavity3d_K2_6": (lambda: cases.cavity(5, 6, dim=3, K=2), 3),
    "cavity3d_6x12": _small_cases()["cavity3d_6x12"],
    "channel_20x8": _small_cases()["channel_20x8"],
    "delaunay_300": _small_cases()["delaunay_300"],
    "batch_4": _small_cases()["batch_4"],
    "agglom_b3": _small_cases()["agglom_b3"],
}


@pytest.mark.parametrize("which", list(_BIT_CASES))
def test_face_values_bit_identical_to_rebuilding(which, monkeypatch):
    """With the face moments on the vector units, face-value mode is the rebuilding mode's
    arithmetic stored and read back: FP64 fields bit-identical at equal cell-kernel shapes.
    Covers the face-value pass's loop for cells whose faces are all internal as well as its
    general loop (the cases mix interior and boundary cells)."""
    monkeypatch.setenv("DUGKS_FACE_MFMA", "0")
    mk, n = _BIT_CASES[which]
    c = mk()
    a, b = Engine(c), Engine(c)
    a.set_face_values(True)
    b.set_face_values(False)
    sa, sb = a.cell_blocks(), b.cell_blocks()
    assert sa == sb, (which, sa, sb)
    a.init_equilibrium()
    b.init_equilibrium()
    for k in (1, n):
        a.step(nsteps=k)
        b.step(nsteps=k)
        ma, mb = a.macro(), b.macro()
        for f in ma:
            assert np.array_equal(ma[f], mb[f]), (which, k, f, rel_l2(ma[f], mb[f]))
    a.close()
    b.close()


@pytest.mark.parametrize("which", list(_small_cases()))
def test_face_moments_on_matrix_cores(which, monkeypatch):
    """FP64 face-value mode sums the face moments of non-wall faces with FP64 MFMA
    (k_face_moments_mx) by default: against the vector-unit sums (DUGKS_FACE_MFMA=0) to
    round-off, and against the oracle no worse than the vector-unit sums (or within 1e-12)."""
    mk, n = _small_cases()[which]
    c = mk()
    monkeypatch.setenv("DUGKS_FACE_MFMA", "1")
    a = Engine(c)
    monkeypatch.setenv("DUGKS_FACE_MFMA", "0")
    b = Engine(c)
    monkeypatch.delenv("DUGKS_FACE_MFMA")
    assert a.face_mfma() and not b.face_mfma()
    o = _oracle(c)
    for e in (a, b):
        e.init_equilibrium()
        e.step(nsteps=n)
    o.step(nsteps=n)
    mo = o.macro()
    for get in ("macro", "face_macro"):
        ma, mb = getattr(a, get)(), getattr(b, get)()
        for k in ma:
            r = rel_l2(ma[k], mb[k])
            _say(f"  {which} {get} {k}: matrix vs vector face moments {r:.2e}")
            # two FP64 summation orders: round-off, up to 1e-11 on fields that nearly cancel
            assert r <= 1e-11, (which, get, k, r)
    for k in mo:
        ra = rel_l2(a.macro()[k], mo[k])
        assert ra <= max(FP64_TOL, 2 * rel_l2(b.macro()[k], mo[k])), (which, k, ra)
    a.close()
    b.close()
    # FP32 storage: the offsets of the stored deviations ride in row 15 of each tile
    monkeypatch.setenv("DUGKS_FACE_MFMA", "1")
    a = Engine(c, storage="fp32")
    monkeypatch.setenv("DUGKS_FACE_MFMA", "0")
    b = Engine(c, storage="fp32")
    monkeypatch.delenv("DUGKS_FACE_MFMA")
    assert a.face_mfma() and not b.face_mfma()
    for e in (a, b):
        e.init_equilibrium()
        e.step(nsteps=n)
    ma, mb = a.macro(), b.macro()
    for k in ma:
        r = rel_l2(ma[k], mb[k])
        _say(f"  {which} fp32 {k}: matrix vs vector face moments {r:.2e}")
        assert r <= 1e-6, (which, "fp32", k, r)
    a.close()
    b.close()


@pytest.mark.parametrize("which", list(_small_cases()))
def test_fused_cell_tail_is_bit_identical(which, monkeypatch):
    """With one workgroup per cell, k_cell forms the cell macros and the next gbar itself
    (k_cell_fin and k_cell_s1 fused into its tail): results bit-identical to the separate
    kernels, FP64 and FP32 storage, over graph-replayed steps."""
    mk, n = _small_cases()[which]
    c = mk()
    for st in ("fp64", "fp32"):
        monkeypatch.setenv("DUGKS_FUSE_S1", "1")
        a = Engine(c, storage=st)
        monkeypatch.setenv("DUGKS_FUSE_S1", "0")
        b = Engine(c, storage=st)
        assert a.fused_cell() and not b.fused_cell(), (a.face_values(), a.cell_blocks())
        for e in (a, b):
            e.init_equilibrium()
            e.step(nsteps=n)
        for get in ("macro", "face_macro"):
            ma, mb = getattr(a, get)(), getattr(b, get)()
            for k in ma:
                assert np.array_equal(ma[k], mb[k]), (which, st, get, k, rel_l2(ma[k], mb[k]))
        cells = np.arange(0, c.n_cells, max(1, c.n_cells // 7), dtype=np.int32)
        ga, gb = a.df(cells), b.df(cells)
        assert np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1]), (which, st, "df")
        _say(f"  {which} {st} fused cell tail: bit-identical")
        a.close()
        b.close()
    monkeypatch.delenv("DUGKS_FUSE_S1")


@pytest.mark.parametrize("which", ["cavity3d_6x12", "cavity2d_12x21", "channel_20x8"])
def test_autotune_shapes_keep_parity(which):
    """The online tuner runs every candidate work decomposition of both kernels on real
    steps; the result still matches the oracle to the FP64 bar, and the shapes it keeps
    tile the padded velocity count."""
    mk, _ = _small_cases()[which]
    c = mk()
    e = Engine(c)
    e.init_equilibrium()
    e.autotune(True)
    n = 24
    e.step(nsteps=n)
    o = _oracle(c)
    o.step(nsteps=n)
    mo, me = o.macro(), e.macro()
    for k in mo:
        assert rel_l2(me[k], mo[k]) <= FP64_TOL, (which, k, rel_l2(me[k], mo[k]))
    nt, nib, nb, nvp = e.blocks()
    ntc, nibc, nbc = e.cell_blocks()
    assert nt * nib * nb == nvp and ntc * nibc * nbc == nvp
    e.close()
    o.close()

