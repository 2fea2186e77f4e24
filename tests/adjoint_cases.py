"""Shared helpers of the adjoint tests: the objective evaluated on the oracle, wall-temperature
perturbations and the small cases that cover the adjoint's code paths."""
from __future__ import annotations

import copy

import numpy as np

from dugksfoam_amd import cases


# boundary kinds whose temperature J depends on
T_PATCHES = (cases.PATCH_WALL, cases.PATCH_FIXED, cases.PATCH_FARFIELD, cases.PATCH_PRESSURE_IN,
             cases.PATCH_PRESSURE_OUT)


def wall_mask(c) -> np.ndarray:
    """Boundary faces with a temperature the adjoint differentiates: walls, fixed, far-field and
    pressure faces."""
    nfi, nb = c.n_internal_faces, c.n_faces - c.n_internal_faces
    m = np.zeros(nb, bool)
    for p in c.patches:
        if p["type"] in T_PATCHES:
            m[p["start"] - nfi: p["start"] - nfi + p["size"]] = True
    return m


def with_wall_T(c, dT):
    """The case with the temperature of boundary face b (walls, fixed, far-field and pressure
    faces) raised by dT[b]."""
    c2 = copy.deepcopy(c)
    nfi = c.n_internal_faces
    for p in c2.patches:
        if p["type"] in T_PATCHES:
            p["T"] = np.asarray(p["T"], float) + dT[p["start"] - nfi: p["start"] - nfi + p["size"]]
    return c2


def objective(backend_cls, c, steps, first, coef) -> float:
    """J = sum_{n >= first} sum_f coef_f rho_f (U_f . S_f) from a backend's face macros."""
    e = backend_cls(c)
    e.init_equilibrium()
    J = 0.0
    for n in range(1, steps + 1):
        e.step(nsteps=1)
        if n >= first:
            fm = e.face_macro()
            J += float(np.sum(coef * fm["rho"] * np.einsum("ij,ij->i", fm["U"], c.Sf)))
    close = getattr(e, "close", None)
    if close:
        close()
    return J


def directional_fd(backend_cls, c, steps, first, coef, v, h=1e-3) -> float:
    """Central difference of J along the wall-temperature direction v (Kelvin per unit step)."""
    Jp = objective(backend_cls, with_wall_T(c, h * v), steps, first, coef)
    Jm = objective(backend_cls, with_wall_T(c, -h * v), steps, first, coef)
    return (Jp - Jm) / (2 * h)


# (name, case builder, steps, first step of the objective, checkpoint spacing)
CASES = [
    ("cavity2d_nv9", lambda: cases.cavity(5, 9), 6, 3, 0),            # xi.S == 0 on axis faces: averaging branches
    ("cavity2d_nv8_every4", lambda: cases.cavity(5, 8), 7, 1, 4),
    ("cavity2d_K2_every1", lambda: cases.cavity(4, 7, K=2), 5, 2, 1),  # h carries K internal degrees of freedom
    ("cavity3d", lambda: cases.cavity(3, 5, dim=3), 4, 2, 0),
    ("channel", lambda: cases.channel(nx=8, ny=4, Lx=1.0, nv=7, kn=0.5, n_modes=3), 6, 1, 2),
    ("channel_hot", lambda: cases.channel(nx=8, ny=4, Lx=1.0, nv=8, kn=0.2, n_modes=3, amp=0.2), 8, 5, 3),
    # fixed-Maxwellian inlet (dJ/dT_b) and zeroGradient outlet
    ("open_zg", lambda: cases.open_channel(6, 4, nv=8, outlet="zeroGradient"), 6, 2, 2),
    # far-field outlet: incoming rho_b M(U of the outlet cell, T_b)
    ("open_far", lambda: cases.open_channel(6, 4, nv=8), 6, 2, 3),
    # pressureIn / pressureOut: rho_b (and the outlet's T_b) from S10 of the previous step (of the
    # initial state before the first step)
    ("pressure", lambda: cases.pressure_channel(6, 4, nv=8), 6, 1, 2),
    # a batch of two cavities with one time step and their own muRef
    ("batch", lambda: cases.union([cases.cavity(4, 6, kn=0.5), cases.cavity(4, 6, kn=0.1)]), 5, 2, 2),
    # symmetry plane: incoming values are the mirrored velocities' outgoing reconstructions
    ("symmetry_nv6", lambda: cases.symmetry_cavity(4, 6, 6), 6, 2, 0),
    ("symmetry_nv7_x0", lambda: cases.symmetry_cavity(3, 6, 7, plane="x0"), 5, 1, 2),  # xi.S == 0: self-mirror
    # symmetryPlane: phi_b = phi_P in the gradient (no lagged row), outgoing values relaxed twice
    ("symplane_nv6", lambda: cases.symmetry_cavity(4, 6, 6, kind="symmetryPlane"), 6, 2, 0),
    ("symplane_nv7_x0", lambda: cases.symmetry_cavity(3, 6, 7, plane="x0", kind="symmetryPlane"), 5, 1, 2),
    # 1-D: heat conduction between two plates at 263 K and 283 K (fvDVM.C:198-215 velocity set)
    ("fourier_1d", lambda: cases.fourier_flow(8, nv=9), 6, 2, 2),
    # polygons with 12 faces, several of them shared with one neighbour
    ("agglomerated", lambda: cases.agglomerated_cavity(9, nv=6, block=3), 5, 2, 2),
]


def random_problem(c, seed=0):
    rng = np.random.default_rng(seed)
    coef = rng.standard_normal(c.n_faces)
    v = rng.standard_normal(c.n_faces - c.n_internal_faces) * wall_mask(c)
    return coef, v
