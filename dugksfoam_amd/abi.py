"""ctypes mirror of include/dugks.h: the case description structs and the
builder that turns a cases.Case into them.

The returned CaseABI object owns every numpy buffer its pointers refer to,
so it must stay alive for the duration of the call that receives it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

DUGKS_ABI_VERSION = 2
STORAGE_FP64 = 0
STORAGE_FP32 = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class DugksPatch(C.Structure):
    _fields_ = [("type", C.c_int32), ("start", C.c_int32), ("size", C.c_int32),
                ("neighbour_patch", C.c_int32), ("U", _dp), ("T", _dp), ("rho", _dp),
                ("pressure", C.c_double)]


class DugksCase(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("dim", C.c_int32), ("n_cells", C.c_int32),
                ("n_internal_faces", C.c_int32), ("n_faces", C.c_int32),
                ("owner", _ip), ("neighbour", _ip), ("C", _dp), ("V", _dp), ("Sf", _dp),
                ("Cf", _dp), ("weights", _dp), ("delta", _dp), ("n_patches", C.c_int32),
                ("patches", C.POINTER(DugksPatch)), ("n_xi", C.c_int32), ("xi", _dp),
                ("xi_weights", _dp), ("xi_max", C.c_double), ("R", C.c_double),
                ("omega", C.c_double), ("T_ref", C.c_double), ("mu_ref", C.c_double),
                ("Pr", C.c_double), ("K", C.c_int32), ("storage", C.c_int32),
                ("device", C.c_int32), ("n_cases", C.c_int32), ("cell_case", _ip),
                ("mu_ref_case", _dp)]


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class CaseABI:
    """Keeps contiguous copies of a Case's arrays and the ctypes struct over them."""

    def __init__(self, case, storage: int = STORAGE_FP64, device: int = 0):
        keep = []

        def d(a, shape=None):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if shape is not None:
                a = a.reshape(shape)
            keep.append(a)
            return dptr(a)

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return iptr(a)

        pats = (DugksPatch * max(1, len(case.patches)))()
        for j, p in enumerate(case.patches):
            pats[j].type = int(p["type"])
            pats[j].start = int(p["start"])
            pats[j].size = int(p["size"])
            pats[j].neighbour_patch = int(p["neighbour_patch"])
            pats[j].U = d(p["U"]) if p.get("U") is not None else _dp()
            pats[j].T = d(p["T"]) if p.get("T") is not None else _dp()
            pats[j].rho = d(p["rho"]) if p.get("rho") is not None else _dp()
            pats[j].pressure = float(p.get("pressure") or 0.0)
        keep.append(pats)
        s = DugksCase()
        s.abi_version = DUGKS_ABI_VERSION
        s.dim = case.dim
        s.n_cells = case.n_cells
        s.n_internal_faces = case.n_internal_faces
        s.n_faces = case.n_faces
        s.owner = i(case.owner)
        s.neighbour = i(case.neighbour)
        s.C = d(case.C)
        s.V = d(case.V)
        s.Sf = d(case.Sf)
        s.Cf = d(case.Cf)
        s.weights = d(case.weights)
        s.delta = d(case.delta if case.delta.size else np.zeros((1, 3)))
        s.n_patches = len(case.patches)
        s.patches = C.cast(pats, C.POINTER(DugksPatch))
        s.n_xi = case.xi.shape[0]
        s.xi = d(case.xi)
        s.xi_weights = d(case.xi_weights)
        s.xi_max = case.xi_max
        s.R = case.R
        s.omega = case.omega
        s.T_ref = case.T_ref
        s.mu_ref = case.mu_ref
        s.Pr = case.Pr
        s.K = case.K
        s.storage = storage
        s.device = device
        if case.n_cases > 1:
            s.n_cases = case.n_cases
            s.cell_case = i(case.cell_case)
            s.mu_ref_case = d(case.mu_ref_case)
        else:
            s.n_cases = 1
            s.cell_case = _ip()
            s.mu_ref_case = _dp()
        self.struct = s
        self._keep = keep

    @property
    def ref(self):
        return C.byref(self.struct)
