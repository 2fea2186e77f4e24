"""ctypes binding of the HIP engine (include/dugks.h, dugksfoam_amd/_build/libdugks.so).

The library is required: every entry point raises if it is missing or fails
to load; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import STORAGE_FP32, STORAGE_FP64, CaseABI, dptr, iptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdugks.so")

# exported symbols of include/dugks.h
SYMBOLS = (
    "dugks_create", "dugks_init_equilibrium", "dugks_step", "dugks_step_cases", "dugks_get_macro",
    "dugks_get_face_macro", "dugks_get_wall", "dugks_get_courant", "dugks_get_df", "dugks_get_df_noalloc", "dugks_get_sizes",
    "dugks_state_size", "dugks_save_state", "dugks_load_state", "dugks_last_step_ms", "dugks_destroy",
    "dugks_last_error", "dugks_abi_version", "dugks_set_profiling", "dugks_get_profile",
    "dugks_get_launch", "dugks_set_graphs", "dugks_set_autotune", "dugks_get_cell_launch",
    "dugks_set_face_values", "dugks_get_face_values", "dugks_autotune_steps_left",
)

_lib = None


class DugksError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dugks error {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH):
    """Loads libdugks.so and declares the C signatures.  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP engine library not built: {path} (run python -m dugksfoam_amd.build)")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.dugks_create.argtypes = [vp, C.POINTER(vp)]
    L.dugks_init_equilibrium.argtypes = [vp, dp, dp, dp]
    L.dugks_step.argtypes = [vp, C.c_double, C.c_int32]
    L.dugks_step_cases.argtypes = [vp, dp, C.c_int32]
    L.dugks_get_macro.argtypes = [vp, dp, dp, dp, dp]
    L.dugks_get_face_macro.argtypes = [vp, dp, dp, dp, dp]
    L.dugks_get_wall.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.dugks_get_courant.argtypes = [vp, C.c_double, dp, dp]
    L.dugks_get_df.argtypes = [vp, C.c_int32, ip, dp, dp]
    L.dugks_get_df_noalloc.argtypes = [vp, C.c_int32, ip, dp, dp]
    L.dugks_get_sizes.argtypes = [vp, ip, ip, ip]
    L.dugks_state_size.argtypes = [vp, C.POINTER(C.c_int64)]
    L.dugks_save_state.argtypes = [vp, vp, C.c_int64]
    L.dugks_load_state.argtypes = [vp, vp, C.c_int64]
    L.dugks_last_step_ms.argtypes = [vp, dp]
    L.dugks_destroy.argtypes = [vp]
    L.dugks_destroy.restype = None
    L.dugks_last_error.argtypes = [vp]
    L.dugks_last_error.restype = C.c_char_p
    L.dugks_abi_version.restype = C.c_int
    L.dugks_set_profiling.argtypes = [vp, C.c_int32]
    L.dugks_get_profile.argtypes = [vp, dp, C.POINTER(C.c_int64)]
    L.dugks_get_launch.argtypes = [vp, ip, ip, ip, ip]
    L.dugks_set_graphs.argtypes = [vp, C.c_int32]
    L.dugks_set_face_values.argtypes = [vp, C.c_int32]
    L.dugks_autotune_steps_left.argtypes = [vp, C.POINTER(C.c_int32)]
    L.dugks_get_face_values.argtypes = [vp, C.POINTER(C.c_int32)]
    L.dugks_set_autotune.argtypes = [vp, C.c_int32]
    L.dugks_get_cell_launch.argtypes = [vp, ip, ip, ip]
    _lib = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _atoi(v: str) -> int:
    """C atoi: optional sign and leading digits, 0 when there are none."""
    import re
    m = re.match(r"\s*([+-]?\d+)", v)
    return int(m.group(1)) if m else 0


class Engine:
    """One DUGKS engine on one GPU for one Case (or a disjoint batch of cases)."""

    def __init__(self, case, storage: str = "fp64", device: int = 0):
        L = load_library()
        self.case = case
        st = {"fp64": STORAGE_FP64, "fp32": STORAGE_FP32}[storage]
        self._abi = CaseABI(case, storage=st, device=device)
        h = C.c_void_p()
        r = L.dugks_create(C.cast(self._abi.ref, C.c_void_p), C.byref(h))
        self._h = h
        if r != 0:
            msg = L.dugks_last_error(h).decode() if h else "no handle"
            if h:
                L.dugks_destroy(h)
            self._h = None
            raise DugksError(r, msg)
        self.storage = storage
        # dugks_create reads DUGKS_FUSE_S1 once (C atoi: unset or nonzero keeps the fusion on)
        v = os.environ.get("DUGKS_FUSE_S1")
        self._fuse_s1 = v is None or _atoi(v) != 0
        v = os.environ.get("DUGKS_FACE_MFMA")
        self._face_mfma = v is None or _atoi(v) != 0
        nv, nc, nf = C.c_int32(), C.c_int32(), C.c_int32()
        L.dugks_get_sizes(h, C.byref(nv), C.byref(nc), C.byref(nf))
        self.nv, self.n_cells, self.n_faces = nv.value, nc.value, nf.value

    def _chk(self, r):
        if r != 0:
            raise DugksError(r, load_library().dugks_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            load_library().dugks_destroy(self._h)
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

    def init_equilibrium(self, rho=None, U=None, T=None):
        c = self.case
        rho = _f64(c.rho0 if rho is None else rho)
        U = _f64(c.U0 if U is None else U)
        T = _f64(c.T0 if T is None else T)
        self._chk(load_library().dugks_init_equilibrium(self._h, dptr(rho), dptr(U), dptr(T)))

    def step(self, dt=None, nsteps: int = 1):
        c = self.case
        if c.n_cases > 1 and dt is None:
            d = _f64(c.dt_case)
            self._chk(load_library().dugks_step_cases(self._h, dptr(d), int(nsteps)))
        else:
            self._chk(load_library().dugks_step(self._h, float(c.dt if dt is None else dt), int(nsteps)))

    def macro(self):
        nc = self.n_cells
        rho, U, T, q = np.zeros(nc), np.zeros((nc, 3)), np.zeros(nc), np.zeros((nc, 3))
        self._chk(load_library().dugks_get_macro(self._h, dptr(rho), dptr(U), dptr(T), dptr(q)))
        return dict(rho=rho, U=U, T=T, q=q)

    def face_macro(self):
        nf = self.n_faces
        rho, U, T, q = np.zeros(nf), np.zeros((nf, 3)), np.zeros(nf), np.zeros((nf, 3))
        self._chk(load_library().dugks_get_face_macro(self._h, dptr(rho), dptr(U), dptr(T), dptr(q)))
        return dict(rho=rho, U=U, T=T, q=q)

    def wall(self, patch: int):
        n = self.case.patches[patch]["size"] if 0 <= patch < len(self.case.patches) else 1
        rw, qw, sw = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 9))
        self._chk(load_library().dugks_get_wall(self._h, int(patch), dptr(rw), dptr(qw), dptr(sw)))
        return dict(rho_w=rw, q_wall=qw, stress_wall=sw)

    def courant(self, dt=None):
        a, b = C.c_double(), C.c_double()
        self._chk(load_library().dugks_get_courant(self._h, float(self.case.dt if dt is None else dt),
                                                   C.byref(a), C.byref(b)))
        return a.value, b.value

    def df(self, cells):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        g = np.zeros((cells.size, self.nv))
        h = np.zeros((cells.size, self.nv))
        self._chk(load_library().dugks_get_df(self._h, cells.size, iptr(cells), dptr(g), dptr(h)))
        return g, h

    def df_noalloc(self, cells):
        """As df(), through dugks_get_df_noalloc (no allocation in the library; the rows pass
        through its fixed per-GPU scratch array)."""
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        g = np.zeros((cells.size, self.nv))
        h = np.zeros((cells.size, self.nv))
        self._chk(load_library().dugks_get_df_noalloc(self._h, cells.size, iptr(cells), dptr(g), dptr(h)))
        return g, h

    def save_state(self) -> bytes:
        n = C.c_int64()
        self._chk(load_library().dugks_state_size(self._h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._chk(load_library().dugks_save_state(self._h, C.cast(buf, C.c_void_p), n.value))
        return buf.raw

    def load_state(self, data: bytes):
        buf = C.create_string_buffer(data, len(data))
        self._chk(load_library().dugks_load_state(self._h, C.cast(buf, C.c_void_p), len(data)))

    def set_profiling(self, on: bool = True):
        """HIP-event timing of the two step kernels on the engine stream."""
        self._chk(load_library().dugks_set_profiling(self._h, 1 if on else 0))

    # kernel slots of dugks_get_profile, in face-value mode and otherwise
    KERNELS_FV = ("k_face_values", "k_face_moments", "k_face_fin", "k_cell", "k_cell_fin", "k_cell_s1")
    KERNELS = ("k_face", None, "k_face_fin", "k_cell", "k_cell_fin", "k_cell_s1")

    def profile(self):
        """({kernel: accumulated ms}, steps) since set_profiling(True)."""
        ms = np.zeros(len(self.KERNELS))
        n = C.c_int64()
        self._chk(load_library().dugks_get_profile(self._h, dptr(ms), C.byref(n)))
        names = self.KERNELS_FV if self.face_values() else self.KERNELS
        out = {k: v for k, v in zip(names, ms.tolist()) if k}
        if self.fused_cell():
            out.pop("k_cell_fin", None)
            out.pop("k_cell_s1", None)
        return out, n.value

    def face_mfma(self) -> bool:
        """True when the face moments of non-wall faces are FP64 matrix products
        (k_face_moments_mx: face-value mode, DUGKS_FACE_MFMA not 0)."""
        return self._face_mfma and self.face_values()

    def fused_cell(self) -> bool:
        """True when k_cell also forms the cell macros and the next step's S1 (face-value mode,
        one workgroup per cell, DUGKS_FUSE_S1 not 0): k_cell_fin and k_cell_s1 do not run."""
        return self._fuse_s1 and self.face_values() and self.cell_blocks()[2] == 1

    def set_graphs(self, on: bool = True):
        """Replay captured two-step HIP graphs for multi-step calls (default on)."""
        self._chk(load_library().dugks_set_graphs(self._h, 1 if on else 0))

    def set_face_values(self, on: bool = True):
        """Face-value mode: stored face values instead of rebuilding them in the cell kernel."""
        self._chk(load_library().dugks_set_face_values(self._h, int(bool(on))))

    def face_values(self) -> bool:
        n = C.c_int32()
        self._chk(load_library().dugks_get_face_values(self._h, C.byref(n)))
        return bool(n.value)

    def blocks(self):
        """Face kernel: (workgroup size, velocity iterations per thread, velocity blocks, padded Nv)."""
        v = [C.c_int32() for _ in range(4)]
        self._chk(load_library().dugks_get_launch(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def cell_blocks(self):
        """Cell kernel: (workgroup size, velocity iterations per thread, velocity blocks)."""
        v = [C.c_int32() for _ in range(3)]
        self._chk(load_library().dugks_get_cell_launch(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, on: bool = True):
        """Time the candidate work decompositions on the next steps and keep the fastest."""
        self._chk(load_library().dugks_set_autotune(self._h, 1 if on else 0))

    def autotune_steps_left(self) -> int:
        n = C.c_int32()
        self._chk(load_library().dugks_autotune_steps_left(self._h, C.byref(n)))
        return n.value

    @property
    def last_step_ms(self) -> float:
        v = C.c_double()
        self._chk(load_library().dugks_last_step_ms(self._h, C.byref(v)))
        return v.value
