"""ctypes binding of the discrete adjoint (the dugks_adjoint_* entry points of include/dugks.h).

Adjoint(case).run(steps, first, coef) returns J and dJ/dT for every boundary face (the wall
temperature of walls, the boundary temperature of fixed, far-field and pressure faces; 0 elsewhere), where

    J = sum over steps n = first .. steps of  sum_f coef[f] (rho U)_f^n . S_f

is a weighted mass flux through faces, (rho U)_f^n the S6 face momentum of step n
(fvDVM.C:473-483).  The run starts from the case's initial fields in equilibrium and advances
FP64 steps of the case's dt; the reverse sweep works from checkpoints.  Like the engine, it needs
the HIP library and a GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import STORAGE_FP64, CaseABI, dptr
from .engine import DugksError, load_library

SYMBOLS = (
    "dugks_adjoint_create", "dugks_adjoint_run", "dugks_adjoint_get_macro", "dugks_adjoint_last_ms",
    "dugks_adjoint_destroy", "dugks_adjoint_last_error",
)


_declared = False


def _lib():
    global _declared
    L = load_library()
    if _declared:
        return L
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.dugks_adjoint_create.argtypes = [vp, C.POINTER(vp)]
    L.dugks_adjoint_run.argtypes = [vp, dp, dp, dp, C.c_double, C.c_int32, C.c_int32, C.c_int32, dp, dp, dp]
    L.dugks_adjoint_get_macro.argtypes = [vp, dp, dp, dp, dp]
    L.dugks_adjoint_last_ms.argtypes = [vp, dp, dp]
    L.dugks_adjoint_destroy.argtypes = [vp]
    L.dugks_adjoint_destroy.restype = None
    L.dugks_adjoint_last_error.argtypes = [vp]
    L.dugks_adjoint_last_error.restype = C.c_char_p
    _declared = True
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Adjoint:
    """Discrete adjoint of the DUGKS step for a case or a batch (every patch kind) on one GPU."""

    def __init__(self, case, device: int = 0):
        L = _lib()
        self.case = case
        self._abi = CaseABI(case, storage=STORAGE_FP64, device=device)
        h = C.c_void_p()
        r = L.dugks_adjoint_create(C.cast(self._abi.ref, C.c_void_p), C.byref(h))
        self._h = h
        if r != 0:
            msg = L.dugks_adjoint_last_error(h).decode() if h else "no handle"
            if h:
                L.dugks_adjoint_destroy(h)
            self._h = None
            raise DugksError(r, msg)

    def _chk(self, r):
        if r != 0:
            raise DugksError(r, _lib().dugks_adjoint_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            _lib().dugks_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, steps: int, first: int, coef, dt=None, checkpoint_every: int = 0, rho=None, U=None, T=None):
        """(J, dJ/dT_w [n_boundary_faces]) of one forward + reverse sweep."""
        c = self.case
        rho = _f64(c.rho0 if rho is None else rho)
        U = _f64(c.U0 if U is None else U)
        T = _f64(c.T0 if T is None else T)
        coef = _f64(coef)
        if coef.shape != (c.n_faces,):
            raise ValueError("coef must hold one value per face")
        J = np.zeros(1)
        g = np.zeros(max(1, c.n_faces - c.n_internal_faces))
        self._chk(_lib().dugks_adjoint_run(self._h, dptr(rho), dptr(U), dptr(T), float(c.dt if dt is None else dt),
                                           int(steps), int(first), int(checkpoint_every), dptr(coef), dptr(J),
                                           dptr(g)))
        return float(J[0]), g[: c.n_faces - c.n_internal_faces]

    def macro(self):
        nc = self.case.n_cells
        rho, U, T, q = np.zeros(nc), np.zeros((nc, 3)), np.zeros(nc), np.zeros((nc, 3))
        self._chk(_lib().dugks_adjoint_get_macro(self._h, dptr(rho), dptr(U), dptr(T), dptr(q)))
        return dict(rho=rho, U=U, T=T, q=q)

    def last_ms(self):
        """(forward sweep ms, reverse sweep ms) of the last run, HIP events."""
        a, b = C.c_double(), C.c_double()
        self._chk(_lib().dugks_adjoint_last_ms(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value
