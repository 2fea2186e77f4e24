"""The dugksFoam application loop over a case directory, with OpenFOAM time-directory output.

    python -m dugksfoam_amd.solver <case-dir> [--steps N] [--fp32] [--no-write]

Solver follows dugksFoam.C:61-108: each iteration takes the Courant number (CourantNo.H,
fvDVM::getCoNum), adapts the time step when adjustTimeStep is on (setDeltaTvar.H:34-49),
advances one DUGKS step (fvDVM::evolution) and, every convergeCheckSteps steps, measures the
relative change of T, rho and U (dugksFoam.C:86-106).  The loop ends at endTime
(Time::run: t < endTime - deltaT/2) or when the change of U falls to convergeTol.  With
writeControl timeStep, rho, U, T and q are written every writeInterval steps as OpenFOAM
ASCII volume fields into <case>/<time>/.

The stepping backend is the HIP engine (dugksfoam_amd.engine.Engine); any object with the
same init_equilibrium / step / macro / courant / wall methods can stand in for it.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from typing import Callable, Optional

import numpy as np

from . import cases, foam

_DIMS = dict(rho="[1 -3 0 0 0 0 0]", U="[0 1 -1 0 0 0 0]", T="[0 0 0 1 0 0 0]", q="[1 0 -3 0 0 0 0]")


def time_name(t: float, precision: int = 6) -> str:
    """Time::timeName with timeFormat general."""
    return f"{t:.{precision}g}"


def _scalar_list(a) -> str:
    return f"{len(a)}\n(\n" + "\n".join(repr(float(v)) for v in a) + "\n)"


def _vector_list(a) -> str:
    return f"{len(a)}\n(\n" + "\n".join("(" + " ".join(repr(float(v)) for v in r) + ")" for r in a) + "\n)"


def _value(a, vector: bool) -> str:
    a = np.asarray(a, dtype=np.float64)
    if a.shape[0] and np.all(a == a[0]):
        v = a[0]
        return "uniform " + ("(" + " ".join(repr(float(x)) for x in v) + ")" if vector else repr(float(v)))
    kind = "vector" if vector else "scalar"
    return f"nonuniform List<{kind}> " + (_vector_list(a) if vector else _scalar_list(a))


def write_field(path: str, name: str, internal, boundary: dict, vector: bool, location: str = ""):
    """An OpenFOAM ASCII volScalarField / volVectorField file.  boundary maps each patch of
    the polyMesh boundary file to (type, face values or None)."""
    cls = "volVectorField" if vector else "volScalarField"
    out = ["FoamFile\n{\n    version     2.0;\n    format      ascii;\n",
           f"    class       {cls};\n    location    \"{location}\";\n    object      {name};\n}}\n\n",
           f"dimensions      {_DIMS.get(name, '[0 0 0 0 0 0 0]')};\n\n",
           f"internalField   {_value(internal, vector)};\n\nboundaryField\n{{\n"]
    for pname, (ptype, val) in boundary.items():
        out.append(f"    {pname}\n    {{\n        type            {ptype};\n")
        if val is not None:
            out.append(f"        value           {_value(val, vector)};\n")
        out.append("    }\n")
    out.append("}\n")
    with open(path, "w") as f:
        f.write("".join(out))


class Solver:
    """dugksFoam's main loop over a case read by foam.read_case."""

    def __init__(self, case_dir: str, storage: str = "fp64", backend=None,
                 log: Optional[Callable[[str], None]] = print):
        self.case_dir = case_dir
        self.case = foam.read_case(case_dir)
        self.ctl = self.case.meta["controls"]
        if backend is None:
            from .engine import Engine
            backend = Engine(self.case, storage=storage)
        self.b = backend
        self.log = log or (lambda s: None)
        self.b.init_equilibrium(self.case.rho0, self.case.U0, self.case.T0)
        self.t = self.ctl["startTime"]
        self.dt = self.ctl["deltaT"]
        self.it = 0
        self.change = dict(T=1.0, rho=1.0, U=1.0)
        # createFields.H:80-82
        self.old = dict(rho=self.case.rho0.copy(), U=self.case.U0.copy(), T=self.case.T0.copy())
        self.macro = None

    def running(self) -> bool:
        """Time::run and the convergence test of dugksFoam.C:63."""
        return self.t < self.ctl["endTime"] - 0.5 * self.dt and self.change["U"] > self.ctl["convergeTol"]

    def iterate(self):
        co, _ = self.b.courant(self.dt)
        if self.ctl["adjustTimeStep"]:
            self.dt = foam.next_delta_t(co, self.dt, self.ctl)
        self.t += self.dt
        self.it += 1
        self.b.step(self.dt, 1)
        self.macro = None
        n = self.ctl["convergeCheckSteps"]
        if self.it % n == 0 and self.it >= n:
            m = self.fields()
            d = {k: np.abs(m[k] - self.old[k]) if k != "U" else np.linalg.norm(m[k] - self.old[k], axis=1)
                 for k in ("T", "rho", "U")}
            self.change = dict(T=float(d["T"].sum() / m["T"].sum()),
                               rho=float(d["rho"].sum() / m["rho"].sum()),
                               U=float(d["U"].sum() / np.linalg.norm(m["U"], axis=1).sum()))
            self.old = dict(rho=m["rho"].copy(), U=m["U"].copy(), T=m["T"].copy())

    def fields(self) -> dict:
        if self.macro is None:
            self.macro = self.b.macro()
        return self.macro

    def write(self):
        """<case>/<time>/{rho,U,T,q}; wall patches carry rho_w, the wall U and T, and qWall;
        cyclic and symmetryPlane patches their type alone; the other non-coupled patches carry
        their boundary cells' values (calculated)."""
        m = self.fields()
        tn = time_name(self.t, self.ctl["timePrecision"])
        d = os.path.join(self.case_dir, tn)
        os.makedirs(d, exist_ok=True)
        live = {p["name"]: i for i, p in enumerate(self.case.patches)}
        bf = {k: {} for k in ("rho", "U", "T", "q")}
        for b in self.case.meta["boundary"]:
            name = b["name"]
            if name not in live:
                for k in bf:
                    bf[k][name] = ("empty", None)
                continue
            p = self.case.patches[live[name]]
            if p["type"] in (cases.PATCH_CYCLIC, cases.PATCH_SYMMETRY_PLANE):
                # OpenFOAM constraint types: written by type alone
                for k in bf:
                    bf[k][name] = ("cyclic" if p["type"] == cases.PATCH_CYCLIC else "symmetryPlane", None)
                continue
            if p["type"] != cases.PATCH_WALL:
                cells = self.case.owner[p["start"]:p["start"] + p["size"]]
                for k in bf:
                    bf[k][name] = ("calculated", m[k][cells])
                continue
            w = self.b.wall(live[name])
            bf["rho"][name] = ("calculatedMaxwell", w["rho_w"])
            bf["U"][name] = ("fixedValue", p["U"])
            bf["T"][name] = ("fixedValue", p["T"])
            bf["q"][name] = ("calculated", w["q_wall"])
        for k, vec in (("rho", False), ("U", True), ("T", False), ("q", True)):
            write_field(os.path.join(d, k), k, m[k], bf[k], vec, location=tn)
        self.write_df()
        return d

    def write_df(self):
        """writeDFonCells (fvDVM.C:877-882, 818-875): for entry i of DFwriteCellList the
        reference writes g~ of cell i -- the list position, not the listed cell id -- for every
        velocity (reference order) to <case>/0/DF<i>, overwriting it at each output time."""
        n = len(self.case.meta.get("df_cells", []))
        if not n:
            return
        g, _ = self.b.df(np.arange(n))
        for i in range(n):
            body = f"{g.shape[1]}\n(\n" + "\n".join(repr(float(v)) for v in g[i]) + "\n)\n"
            with open(os.path.join(self.case_dir, "0", f"DF{i}"), "w") as f:
                f.write("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
                        f"    class       scalarList;\n    location    \"0\";\n    object      DF{i};\n}}\n\n" + body)

    def run(self, max_steps: Optional[int] = None, write: bool = True) -> int:
        every = int(self.ctl["writeInterval"]) if self.ctl["writeControl"] == "timeStep" else 0
        t0 = time.time()
        while self.running() and (max_steps is None or self.it < max_steps):
            self.iterate()
            if write and every and self.it % every == 0:
                self.write()
            if self.it % 100 == 0:
                self.log(f"Step = {self.it}  Time = {self.t:.6g}  deltaT = {self.dt:.6g}  "
                         f"Uchange = {self.change['U']:.3e}  ClockTime = {time.time() - t0:.1f} s")
        if write and (not every or self.it % every):
            self.write()
        return self.it


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("case")
    ap.add_argument("--steps", type=int, default=None, help="stop after this many steps")
    ap.add_argument("--fp32", action="store_true", help="FP32 distribution storage")
    ap.add_argument("--no-write", action="store_true")
    a = ap.parse_args(argv)
    s = Solver(a.case, storage="fp32" if a.fp32 else "fp64")
    n = s.run(max_steps=a.steps, write=not a.no_write)
    print(f"{n} steps, t = {s.t:.9g}, deltaT = {s.dt:.6g}, Uchange = {s.change['U']:.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
