"""ctypes loader of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from dugksfoam_amd.abi import CaseABI, dptr, iptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
# Same source built with x87 extended-precision intermediates: its distance from the FP64
# build measures the oracle's own round-off (the floor below which parity is not resolvable).
LIB_EXT = os.path.join(HERE, "_build", "liboracle_ext.so")
# Same source with every velocity pass reversed: sensitivity to summation order.
LIB_REV = os.path.join(HERE, "_build", "liboracle_rev.so")
# Same source built with AddressSanitizer and UndefinedBehaviorSanitizer (make san); loadable only
# into a process started with the compiler's libasan preloaded.
LIB_SAN = os.path.join(HERE, "_build", "liboracle_san.so")
_libs = {}


def build(force: bool = False) -> str:
    """Compiles oracle.c with its Makefile (gcc, -ffp-contract=off, OpenMP), both builds."""
    cmd = ["make", "-s", "-C", HERE]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return LIB


def lib(variant: str = "fp64"):
    path = {"fp64": LIB, "ext": LIB_EXT, "rev": LIB_REV, "san": LIB_SAN}[variant]
    if variant not in _libs:
        if variant == "san":
            subprocess.run(["make", "-s", "-C", HERE, "san"], check=True)
        elif not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [vp]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_threads.argtypes = [vp, C.c_int]
        L.oracle_set_cache_faces.argtypes = [vp, C.c_int]
        L.oracle_init_equilibrium.argtypes = [vp, dp, dp, dp]
        L.oracle_step.argtypes = [vp, C.c_double, C.c_int]
        L.oracle_step_cases.argtypes = [vp, dp, C.c_int]
        L.oracle_get_macro.argtypes = [vp, dp, dp, dp, dp]
        L.oracle_get_face_macro.argtypes = [vp, dp, dp, dp, dp]
        L.oracle_get_wall.argtypes = [vp, C.c_int, dp, dp, dp]
        L.oracle_get_df.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), dp, dp]
        L.oracle_get_courant.argtypes = [vp, C.c_double, dp, dp]
        L.oracle_lsq_vectors.argtypes = [vp, dp, dp]
        L.oracle_lsq_grad.argtypes = [vp, dp, dp, dp]
        L.oracle_shakhov.argtypes = [vp, C.c_double, dp, C.c_double, dp, dp, dp, dp]
        L.oracle_maxwell_by_rho.restype = C.c_double
        L.oracle_maxwell_by_rho.argtypes = [vp, dp, C.c_double, dp]
        L.oracle_tau.restype = C.c_double
        L.oracle_tau.argtypes = [vp, C.c_double, C.c_double]
        L.oracle_velocities.argtypes = [vp, dp, dp]
        L.oracle_incoming.argtypes = [vp, dp]
        L.oracle_boundary_state.argtypes = [vp, dp, dp]
        L.oracle_total_mass.restype = C.c_double
        L.oracle_total_mass.argtypes = [vp]
        _libs[variant] = L
    return _libs[variant]


def _a(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Oracle:
    """FP64 reference solver for one Case."""

    def __init__(self, case, threads: int = 1, cache_faces: bool = False, variant: str = "fp64"):
        """variant "ext": extended-precision intermediates; "rev": velocity passes in reverse
        order (both estimate the round-off floor of a comparison)."""
        self.case = case
        self._abi = CaseABI(case)
        self._L = L = lib(variant)
        self._h = L.oracle_create(C.cast(self._abi.ref, C.c_void_p))
        if not self._h:
            raise RuntimeError("oracle_create failed")
        L.oracle_set_threads(self._h, int(threads))
        if cache_faces and L.oracle_set_cache_faces(self._h, 1) != 0:
            raise MemoryError("oracle face cache")
        self.nv = case.nv

    def close(self):
        if self._h:
            self._L.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_equilibrium(self, rho=None, U=None, T=None):
        c = self.case
        rho = _a(c.rho0 if rho is None else rho)
        U = _a(c.U0 if U is None else U)
        T = _a(c.T0 if T is None else T)
        r = self._L.oracle_init_equilibrium(self._h, dptr(rho), dptr(U), dptr(T))
        if r:
            raise RuntimeError(f"oracle_init_equilibrium {r}")

    def step(self, dt=None, nsteps: int = 1):
        c = self.case
        if c.n_cases > 1 and dt is None:
            dtc = _a(c.dt_case)
            r = self._L.oracle_step_cases(self._h, dptr(dtc), int(nsteps))
        else:
            r = self._L.oracle_step(self._h, float(c.dt if dt is None else dt), int(nsteps))
        if r:
            raise RuntimeError(f"oracle_step {r}")

    def macro(self):
        nc = self.case.n_cells
        rho, U, T, q = np.zeros(nc), np.zeros((nc, 3)), np.zeros(nc), np.zeros((nc, 3))
        self._L.oracle_get_macro(self._h, dptr(rho), dptr(U), dptr(T), dptr(q))
        return dict(rho=rho, U=U, T=T, q=q)

    def face_macro(self):
        nf = self.case.n_faces
        rho, U, T, q = np.zeros(nf), np.zeros((nf, 3)), np.zeros(nf), np.zeros((nf, 3))
        self._L.oracle_get_face_macro(self._h, dptr(rho), dptr(U), dptr(T), dptr(q))
        return dict(rho=rho, U=U, T=T, q=q)

    def wall(self, patch: int):
        n = self.case.patches[patch]["size"]
        rw, qw, sw = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 9))
        r = self._L.oracle_get_wall(self._h, int(patch), dptr(rw), dptr(qw), dptr(sw))
        if r:
            raise RuntimeError("oracle_get_wall")
        return dict(rho_w=rw, q_wall=qw, stress_wall=sw)

    def df(self, cells):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        g = np.zeros((cells.size, self.nv))
        h = np.zeros((cells.size, self.nv))
        self._L.oracle_get_df(self._h, cells.size, iptr(cells), dptr(g), dptr(h))
        return g, h

    def courant(self, dt=None):
        a, b = C.c_double(), C.c_double()
        self._L.oracle_get_courant(self._h, float(self.case.dt if dt is None else dt), C.byref(a), C.byref(b))
        return a.value, b.value

    def lsq_vectors(self):
        p = np.zeros((self.case.n_faces, 3))
        n = np.zeros((max(1, self.case.n_internal_faces), 3))
        self._L.oracle_lsq_vectors(self._h, dptr(p), dptr(n))
        return p, n[: self.case.n_internal_faces]

    def lsq_grad(self, phi, phib):
        g = np.zeros((self.case.n_cells, 3))
        phi, phib = _a(phi), _a(phib)
        if phib.size == 0:
            phib = np.zeros(1)
        self._L.oracle_lsq_grad(self._h, dptr(phi), dptr(phib), dptr(g))
        return g

    def shakhov(self, rho, U, T, q, xi):
        g, h = C.c_double(), C.c_double()
        self._L.oracle_shakhov(self._h, float(rho), dptr(_a(U)), float(T), dptr(_a(q)), dptr(_a(xi)),
                             C.byref(g), C.byref(h))
        return g.value, h.value

    def maxwell_by_rho(self, U, T, xi):
        return self._L.oracle_maxwell_by_rho(self._h, dptr(_a(U)), float(T), dptr(_a(xi)))

    def tau(self, T, rho):
        return self._L.oracle_tau(self._h, float(T), float(rho))

    def velocities(self):
        xi = np.zeros((self.nv, 3))
        w = np.zeros(self.nv)
        self._L.oracle_velocities(self._h, dptr(xi), dptr(w))
        return xi, w

    def incoming(self):
        nb = self.case.n_faces - self.case.n_internal_faces
        a = np.zeros(max(1, nb))
        self._L.oracle_incoming(self._h, dptr(a))
        return a[:nb]

    def boundary_state(self):
        """rho_b, T_b of every boundary face (pressure patches: as the last S10 set them)."""
        nb = self.case.n_faces - self.case.n_internal_faces
        rb, Tb = np.zeros(max(1, nb)), np.zeros(max(1, nb))
        self._L.oracle_boundary_state(self._h, dptr(rb), dptr(Tb))
        return rb[:nb], Tb[:nb]

    def total_mass(self):
        return self._L.oracle_total_mass(self._h)
