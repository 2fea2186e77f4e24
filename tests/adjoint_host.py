"""Builds and binds tests/adjoint_host.cpp: adjoint_core.h and adjoint_driver.h run on the host.

Test infrastructure (see adjoint_host.cpp).  HostAdjoint mirrors dugksfoam_amd.adjoint.Adjoint.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from dugksfoam_amd.abi import STORAGE_FP64, CaseABI, dptr

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "adjoint_host.cpp")
OUT = os.path.join(HERE, "_build", "libadjoint_host.so")
DEPS = [SRC] + [os.path.join(ROOT, "dugksfoam_amd", "csrc", n) for n in ("adjoint_core.h", "adjoint_driver.h", "lsq.h")]

_lib = None


# AddressSanitizer + UndefinedBehaviorSanitizer build (CPU test suite only; loadable into a process
# started with the compiler's libasan preloaded), chosen with DUGKS_ADJOINT_HOST_SAN=1
OUT_SAN = os.path.join(HERE, "_build", "libadjoint_host_san.so")
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def build(force: bool = False, sanitize: bool = False) -> str:
    out = OUT_SAN if sanitize else OUT
    fresh = os.path.exists(out) and all(os.path.getmtime(out) >= os.path.getmtime(d) for d in DEPS)
    if force or not fresh:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        tmp = out + ".tmp"
        opt = SAN_FLAGS if sanitize else ["-O2"]
        subprocess.run(["g++", *opt, "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
                        "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "include"), SRC, "-o", tmp], check=True)
        os.replace(tmp, out)
    return out


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(build(sanitize=os.environ.get("DUGKS_ADJOINT_HOST_SAN") == "1"))
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.adjoint_host_create.argtypes = [vp, C.POINTER(vp)]
        L.adjoint_host_run.argtypes = [vp, dp, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.adjoint_host_macro.argtypes = [vp, dp, dp, dp, dp]
        L.adjoint_host_error.argtypes = [vp]
        L.adjoint_host_error.restype = C.c_char_p
        L.adjoint_host_destroy.argtypes = [vp]
        L.adjoint_host_destroy.restype = None
        _lib = L
    return _lib


class HostAdjoint:
    def __init__(self, case):
        L = lib()
        self.case = case
        self._abi = CaseABI(case, storage=STORAGE_FP64, device=0)
        h = C.c_void_p()
        r = L.adjoint_host_create(C.cast(self._abi.ref, C.c_void_p), C.byref(h))
        self._h = h
        if r != 0:
            msg = L.adjoint_host_error(h).decode()
            L.adjoint_host_destroy(h)
            self._h = None
            raise RuntimeError(f"adjoint_host_create {r}: {msg}")

    def run(self, steps, first, coef, dt=None, checkpoint_every=0):
        c = self.case
        rho, U, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (c.rho0, c.U0, c.T0))
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        J = np.zeros(1)
        g = np.zeros(max(1, c.n_faces - c.n_internal_faces))
        r = lib().adjoint_host_run(self._h, dptr(rho), dptr(U), dptr(T), float(c.dt if dt is None else dt),
                                   int(steps), int(first), int(checkpoint_every), dptr(coef), dptr(J), dptr(g))
        if r != 0:
            raise RuntimeError(f"adjoint_host_run {r}: {lib().adjoint_host_error(self._h).decode()}")
        return float(J[0]), g[: c.n_faces - c.n_internal_faces]

    def macro(self):
        nc = self.case.n_cells
        rho, U, T, q = np.zeros(nc), np.zeros((nc, 3)), np.zeros(nc), np.zeros((nc, 3))
        r = lib().adjoint_host_macro(self._h, dptr(rho), dptr(U), dptr(T), dptr(q))
        if r != 0:
            raise RuntimeError(lib().adjoint_host_error(self._h).decode())
        return dict(rho=rho, U=U, T=T, q=q)

    def close(self):
        if getattr(self, "_h", None):
            lib().adjoint_host_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
