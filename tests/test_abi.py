"""The C-ABI library: it loads, exports every function include/dugks.h declares,
its case struct matches the ctypes mirror, and it rejects use without a GPU
cleanly.  No compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT
from dugksfoam_amd import abi, adjoint, cases, engine

HEADER = os.path.join(ROOT, "include", "dugks.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\s*\*)\s*(dugks_\w+)\s*\(", txt, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(engine.LIB_PATH):
        from dugksfoam_amd import build
        build.build()
    return engine.load_library()


def test_library_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 17
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (dugks_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(engine.SYMBOLS) | set(adjoint.SYMBOLS) == set(names)
    for n in names:
        assert hasattr(lib, n)


def test_abi_version(lib):
    assert lib.dugks_abi_version() == abi.DUGKS_ABI_VERSION == 2


def test_struct_layout_matches_header(tmp_path):
    """offsetof of every dugks_case / dugks_patch field as the C compiler lays it out."""
    fields = [f[0] for f in abi.DugksCase._fields_]
    pfields = [f[0] for f in abi.DugksPatch._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for f in fields:
        src.append(f'printf("case {f} %zu\\n", offsetof(dugks_case, {f}));')
    for f in pfields:
        src.append(f'printf("patch {f} %zu\\n", offsetof(dugks_patch, {f}));')
    src.append('printf("size case %zu patch %zu\\n", sizeof(dugks_case), sizeof(dugks_patch)); return 0;}')
    cfile = tmp_path / "layout.c"
    cfile.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(cfile), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        p = line.split()
        if len(p) == 3:
            got[(p[0], p[1])] = int(p[2])
    for f in fields:
        assert got[("case", f)] == getattr(abi.DugksCase, f).offset, f
    for f in pfields:
        assert got[("patch", f)] == getattr(abi.DugksPatch, f).offset, f
    sz = [l for l in out if l.startswith("size")][0].split()
    assert int(sz[2]) == C.sizeof(abi.DugksCase) and int(sz[4]) == C.sizeof(abi.DugksPatch)


def test_create_validates_arguments(lib):
    """Null handle arguments and a wrong ABI version are rejected before any device use."""
    h = C.c_void_p()
    assert lib.dugks_create(None, C.byref(h)) == -1
    c = cases.cavity(3, 5)
    a = abi.CaseABI(c)
    a.struct.abi_version = 99
    r = lib.dugks_create(C.cast(a.ref, C.c_void_p), C.byref(h))
    assert r == -1
    assert b"abi_version" in lib.dugks_last_error(h)
    lib.dugks_destroy(h)
    a.struct.abi_version = abi.DUGKS_ABI_VERSION
    a.struct.dim = 4
    r = lib.dugks_create(C.cast(a.ref, C.c_void_p), C.byref(h))
    assert r == -6
    lib.dugks_destroy(h)


def test_no_cpu_fallback_without_gpu(lib):
    """Without a visible GPU the engine refuses to run (no silent CPU path)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(engine.DugksError) as e:
        engine.Engine(cases.cavity(3, 5))
    assert e.value.code == -2


def test_adjoint_no_cpu_fallback_without_gpu(lib):
    """The adjoint, too, refuses to run without a GPU, and explains why."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(engine.DugksError) as e:
        adjoint.Adjoint(cases.cavity(3, 5))
    assert e.value.code in (-1, -2) and "HIP device" in str(e.value)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        engine.load_library(str(tmp_path / "nope.so"))
