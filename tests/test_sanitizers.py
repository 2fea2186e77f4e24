"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer: the FP64 oracle
(oracle/oracle.c) and the host build of the discrete adjoint (tests/adjoint_host.cpp with
adjoint_core.h, adjoint_driver.h and lsq.h), each built with -fsanitize=address,undefined and
-fno-sanitize-recover, run over small cases of every boundary kind in a child Python process
that preloads the compiler's libasan.  Any out-of-bounds access, use after free, overflow or
other undefined behaviour aborts the child, and the test fails with the sanitizer's report."""
import os
import subprocess
import sys
import textwrap

import pytest

from conftest import ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def _libasan():
    try:
        p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    p = p.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _run_sanitized(script: str, extra_env=None, timeout=600):
    asan = _libasan()
    if asan is None:
        pytest.skip("gcc's libasan is not available")
    env = dict(os.environ)
    env.update({"LD_PRELOAD": asan,
                # the interpreter itself is not instrumented: its allocations are not leaks of ours
                "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=86",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1:exitcode=87",
                "PYTHONPATH": os.pathsep.join([ROOT, HERE])})
    env.update(extra_env or {})
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(script)], env=env, capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    report = (r.stdout + r.stderr)[-6000:]
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in report and "runtime error" not in report, report
    return r.stdout


def test_oracle_under_sanitizers():
    out = _run_sanitized("""
        import numpy as np
        from dugksfoam_amd import cases
        from oracle.oracle import Oracle
        cs = [cases.cavity(5, 9), cases.cavity(3, 6, dim=3, kn=0.1), cases.cavity(3, 5, dim=3, K=2),
              cases.channel(6, 4, 1.0, 0.6, nv=8), cases.symmetry_cavity(4, 6, nv=9),
              cases.symmetry_cavity(3, 6, nv=8, plane="x0"), cases.open_channel(6, 4, nv=8),
              cases.pressure_channel(6, 4, nv=8), cases.open_channel(6, 4, nv=8, outlet="zeroGradient"),
              cases.delaunay_cavity(60, 6, nv=6), cases.batch(n_cases=3, nmin=3, nmax=5, nv=6, seed=2),
              cases.symmetry_cavity(4, 6, nv=9, kind="symmetryPlane"), cases.fourier_flow(8, nv=9),
              cases.agglomerated_cavity(9, nv=6, block=3), cases.delaunay_cavity(60, 6, nv=6, wall_layer=True)]
        for c in cs:
            for threads in (1, 3):
                o = Oracle(c, threads=threads, variant="san", cache_faces=threads == 3)
                o.init_equilibrium()
                o.step(nsteps=3)
                m, f = o.macro(), o.face_macro()
                assert all(np.all(np.isfinite(v)) for v in list(m.values()) + list(f.values()))
                for p, pp in enumerate(c.patches):
                    if pp["type"] == 1:
                        o.wall(p)
                o.df(np.arange(min(3, c.n_cells)))
                o.courant()
                o.lsq_vectors()
                o.boundary_state()
                o.close()
            print("ok", c.name, flush=True)
    """)
    assert out.count("ok") == 15


def test_host_adjoint_under_sanitizers():
    out = _run_sanitized("""
        import numpy as np
        from adjoint_cases import CASES
        from adjoint_host import HostAdjoint
        for name, mk, steps, first, every in CASES:
            c = mk()
            coef = np.random.default_rng(0).standard_normal(c.n_faces)
            A = HostAdjoint(c)
            J, g = A.run(steps, first, coef, checkpoint_every=every)
            assert np.isfinite(J) and np.all(np.isfinite(g))
            A.macro()
            A.close()
            print("ok", name, flush=True)
    """, extra_env={"DUGKS_ADJOINT_HOST_SAN": "1"})
    assert out.count("ok") >= 10
