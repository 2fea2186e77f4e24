"""Mass-flow objective of the periodic micro-channel and its gradient with respect to the
lower-wall temperature coefficients.

The workload: cases.channel (200 x 40 cells, 32^2 velocities, cyclic in x, lower wall at
T(x) = T0 (1 + sum_k a_k sin(2 pi k x / L)), upper wall at T0), run for N steps from rest; the
objective J(a) is the mass flow through the plane x = L/2 averaged over the last M steps,

    J = (1/M) sum_{n > N - M} sum_{f on the plane} rho_f (U_f . S_f),

with rho_f, U_f the face macros of S6 (fvDVM.C:473-495, dugks_get_face_macro).

Two ways to the gradient dJ/da_k:
  adjoint_gradient  the discrete adjoint (dugksfoam_amd.adjoint): one forward and one reverse
                    sweep give dJ/dT_w on every wall face, chained through
                    dT_w/da_k = T0 sin(2 pi k x_f / L) on the lower wall;
  gradient          central differences of the FP64 engine, (J(a + h e_k) - J(a - h e_k)) / 2h,
                    i.e. 2 n_modes forward runs.
mass_flow_function wraps J in a torch.autograd.Function whose backward returns either.
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import numpy as np

from . import cases


def plane_faces(case: cases.Case, x: float, axis: int = 0) -> np.ndarray:
    """Internal faces lying on the plane x_axis = x with their normal along that axis."""
    nfi = case.n_internal_faces
    cf = case.Cf[:nfi, axis]
    sf = case.Sf[:nfi]
    on = np.abs(cf - x) <= 1e-9 * max(1.0, abs(x))
    along = np.abs(sf[:, axis]) > 0.5 * np.linalg.norm(sf, axis=1)
    return np.nonzero(on & along)[0]


def mass_flow(case: cases.Case, steps: int, avg_last: int, backend: Optional[Callable] = None) -> float:
    """J for one case: steps - avg_last steps, then the plane mass flow after each of the last
    avg_last steps, averaged.  backend(case) returns an object with the Engine methods
    init_equilibrium / step / face_macro / close (default: the HIP engine, FP64)."""
    if backend is None:
        from .engine import Engine
        backend = Engine
    faces = plane_faces(case, 0.5 * case.meta["Lx"])
    if faces.size == 0:
        raise ValueError("no faces on the mid plane")
    S = case.Sf[faces]
    e = backend(case)
    try:
        e.init_equilibrium()
        if steps > avg_last:
            e.step(nsteps=steps - avg_last)
        acc = 0.0
        for _ in range(avg_last):
            e.step(nsteps=1)
            fm = e.face_macro()
            acc += float(np.sum(fm["rho"][faces] * np.einsum("ij,ij->i", fm["U"][faces], S)))
        return acc / avg_last
    finally:
        close = getattr(e, "close", None)
        if close:
            close()


def gradient(a: Sequence[float], steps: int = 2000, avg_last: int = 500, h: float = 1e-6,
             backend: Optional[Callable] = None, **channel_kw):
    """(J(a), dJ/da) for the channel with coefficients a (central differences, step h)."""
    a = np.asarray(a, dtype=np.float64)
    J = mass_flow(cases.channel(a=a, n_modes=a.size, **channel_kw), steps, avg_last, backend)
    g = np.zeros_like(a)
    for k in range(a.size):
        ap, am = a.copy(), a.copy()
        ap[k] += h
        am[k] -= h
        Jp = mass_flow(cases.channel(a=ap, n_modes=a.size, **channel_kw), steps, avg_last, backend)
        Jm = mass_flow(cases.channel(a=am, n_modes=a.size, **channel_kw), steps, avg_last, backend)
        g[k] = (Jp - Jm) / (2.0 * h)
    return J, g


def objective_coef(case: cases.Case, avg_last: int) -> np.ndarray:
    """Per-face weights c_f of the mid-plane mass flow averaged over avg_last steps."""
    faces = plane_faces(case, 0.5 * case.meta["Lx"])
    if faces.size == 0:
        raise ValueError("no faces on the mid plane")
    coef = np.zeros(case.n_faces)
    coef[faces] = 1.0 / avg_last
    return coef


def dTw_da(case: cases.Case) -> tuple:
    """(boundary-face indices b of the lower wall, dT_w(b)/da_k [n_b, n_modes])."""
    nfi = case.n_internal_faces
    low = next(p for p in case.patches if p["name"] == "lowerWall")
    b = np.arange(low["start"], low["start"] + low["size"]) - nfi
    xb = case.Cf[nfi + b, 0]
    k = np.arange(1, np.asarray(case.meta["a"]).size + 1)
    return b, case.meta["T0"] * np.sin(2 * np.pi * k[None, :] * xb[:, None] / case.meta["Lx"])


def adjoint_gradient(a: Sequence[float], steps: int = 2000, avg_last: int = 500, backend: Optional[Callable] = None,
                     checkpoint_every: int = 100, **channel_kw):
    """(J(a), dJ/da) from one forward and one reverse sweep of the discrete adjoint.
    backend(case) returns an object with the dugksfoam_amd.adjoint.Adjoint run method."""
    if backend is None:
        from .adjoint import Adjoint
        backend = Adjoint
    a = np.asarray(a, dtype=np.float64)
    c = cases.channel(a=a, n_modes=a.size, **channel_kw)
    if not 1 <= avg_last <= steps:
        raise ValueError("need 1 <= avg_last <= steps")
    adj = backend(c)
    try:
        J, g = adj.run(steps, steps - avg_last + 1, objective_coef(c, avg_last),
                       checkpoint_every=min(checkpoint_every, steps))
    finally:
        close = getattr(adj, "close", None)
        if close:
            close()
    b, dT = dTw_da(c)
    return J, dT.T @ g[b]


def mass_flow_function(steps: int = 2000, avg_last: int = 500, h: float = 1e-6, method: str = "adjoint",
                       backend: Optional[Callable] = None, **channel_kw):
    """A torch.autograd.Function computing J(a) (forward) and dJ/da (backward), the gradient
    from the discrete adjoint (method "adjoint"; backend(case) as for adjoint_gradient) or central
    differences (method "fd"; backend(case) as for mass_flow).  J and the gradient are returned
    on the device and in the dtype of the coefficients a.  With a on the GPU, torch's device
    must be initialised before the engine library first loads (torch.ones(1, device="cuda")),
    so that both use torch's HIP runtime: a process whose first HIP user is the engine library
    keeps torch from seeing the device."""
    import torch

    if method not in ("adjoint", "fd"):
        raise ValueError("method must be 'adjoint' or 'fd'")

    class MassFlow(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a):
            an = a.detach().cpu().numpy().astype(np.float64)
            if method == "adjoint":
                J, g = adjoint_gradient(an, steps, avg_last, backend=backend, **channel_kw)
            else:
                J, g = gradient(an, steps, avg_last, h, backend=backend, **channel_kw)
            ctx.save_for_backward(torch.as_tensor(g, dtype=a.dtype, device=a.device))
            return torch.as_tensor(J, dtype=a.dtype, device=a.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g

    return MassFlow
