"""Case front-end for dugksFoam case directories (OpenFOAM ASCII format).

read_case(path) builds a Case (dugksfoam_amd.cases.Case) from the files a dugksFoam run reads:

  constant/polyMesh/{points,faces,owner,neighbour,boundary}   the mesh (polyMesh)
  0/{rho,U,T}                     initial fields (uniform or nonuniform internalField) and the
                                  boundary types: rho calculatedMaxwell + U, T fixedValue is a
                                  diffuse Maxwell wall (doc/usage.tex:61-75); rho symmetryMod is a
                                  mirror plane, and so is a symmetryPlane patch (mesh and fields,
                                  doc/usage.tex:77-81); rho farField / fixedValue with T fixedValue are
                                  open boundaries with an incoming Maxwellian; empty patches are
                                  omitted (fvPatch size 0); cyclic patches are coupled pairs
  constant/DVMProperties          fvDVMparas {nDV, xiMax, xiMin}, gasProperties
                                  {R, omega, Tref, muRef, Pr, KInner = 0} (fvDVM.C:913-933),
                                  DFwriteCellList (fvDVM.C:1067)
  constant/Xis, constant/weights  the 1-D discrete-velocity set (fvDVM.C:56-117)
  system/controlDict              deltaT, adjustTimeStep, maxCo

Geometry follows OpenFOAM's primitiveMesh: face centres and area vectors by triangle
decomposition about the point average of the face; cell centres and volumes by pyramid
decomposition about the average of the cell's face centres.  Interpolation weights and
boundary deltas are those of cases._assemble (surfaceInterpolation::makeWeights, the
patch-normal fvPatch::delta).  Patch types other than walls, symmetryMod and symmetryPlane planes,
farField, fixedValue, pressureIn/pressureOut and zeroGradient open boundaries, cyclics and empty
are rejected.
"""
from __future__ import annotations

import math
import os
import re
from typing import Dict, List, Tuple

import numpy as np

from . import cases

_COMMENT = re.compile(r"//[^\n]*|/\*.*?\*/", re.S)
VSMALL = 1.0e-300


def _strip(text: str) -> str:
    """Removes comments and the FoamFile header dictionary."""
    text = _COMMENT.sub(" ", text)
    m = re.search(r"\bFoamFile\s*\{", text)
    if m:
        depth, i = 1, m.end()
        while depth and i < len(text):
            depth += {"{": 1, "}": -1}.get(text[i], 0)
            i += 1
        text = text[:m.start()] + text[i:]
    return text


def _read(path: str) -> str:
    with open(path) as f:
        return _strip(f.read())


def _list_body(text: str) -> Tuple[int, str]:
    """Count and body of the first top-level list `N ( ... )`."""
    m = re.search(r"(\d+)\s*\(", text)
    if not m:
        raise ValueError("no list found")
    n = int(m.group(1))
    depth, i = 1, m.end()
    start = i
    while depth and i < len(text):
        c = text[i]
        if c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
        i += 1
    return n, text[start:i - 1]


def read_points(path: str) -> np.ndarray:
    n, body = _list_body(_read(path))
    pts = np.array(re.findall(r"\(([^()]*)\)", body), dtype=object)
    p = np.array([[float(v) for v in s.split()] for s in pts], dtype=np.float64)
    if p.shape != (n, 3):
        raise ValueError(f"{path}: expected {n} points")
    return p


def read_faces(path: str) -> List[np.ndarray]:
    n, body = _list_body(_read(path))
    faces = [np.array(s.split(), dtype=np.int64) for s in re.findall(r"\d+\s*\(([^()]*)\)", body)]
    if len(faces) != n:
        raise ValueError(f"{path}: expected {n} faces")
    return faces


def read_labels(path: str) -> np.ndarray:
    n, body = _list_body(_read(path))
    a = np.array(body.split(), dtype=np.int64)
    if a.size != n:
        raise ValueError(f"{path}: expected {n} labels")
    return a


def _dict_entries(text: str) -> Dict[str, str]:
    """Top-level `name { ... }` sub-dictionaries of a dictionary body."""
    out, i = {}, 0
    for m in re.finditer(r"(\w+)\s*\{", text):
        if m.start() < i:
            continue
        depth, j = 1, m.end()
        while depth and j < len(text):
            depth += {"{": 1, "}": -1}.get(text[j], 0)
            j += 1
        out[m.group(1)] = text[m.end():j - 1]
        i = j
    return out


def _keyword(body: str, key: str):
    m = re.search(rf"(?<![\w.]){key}\s+([^;]*);", body)
    return None if m is None else m.group(1).strip()


def _number(s: str) -> float:
    """Last number of an entry (plain or dimensioned `name [dims] value`)."""
    return float(re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", s)[-1])


def read_boundary(path: str) -> List[dict]:
    _, body = _list_body(_read(path))
    patches = []
    for name, b in _dict_entries(body).items():
        patches.append(dict(name=name, type=_keyword(b, "type"), nFaces=int(_keyword(b, "nFaces")),
                            startFace=int(_keyword(b, "startFace")),
                            neighbourPatch=_keyword(b, "neighbourPatch")))
    return patches


def _field_value(spec: str, n: int, vector: bool) -> np.ndarray:
    """`uniform v`, `uniform (x y z)` or `nonuniform List<...> N (...)` as an [n] / [n,3] array."""
    spec = spec.strip()
    if spec.startswith("uniform"):
        vals = [float(v) for v in re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", spec[7:])]
        return np.tile(np.array(vals), (n, 1)) if vector else np.full(n, vals[0])
    if spec.startswith("nonuniform"):
        cnt, body = _list_body(spec)
        if vector:
            a = np.array([[float(v) for v in s.split()] for s in re.findall(r"\(([^()]*)\)", body)])
        else:
            a = np.array(body.split(), dtype=np.float64)
        if a.shape[0] != n:
            raise ValueError("nonuniform field size mismatch")
        return a
    raise ValueError(f"unsupported field value: {spec[:40]}")


def read_field(path: str, n_cells: int, patch_sizes: Dict[str, int], vector: bool):
    text = _read(path)
    m = re.search(r"internalField\s+(.*?);", text, re.S)
    internal = _field_value(m.group(1), n_cells, vector)
    bm = re.search(r"boundaryField\s*\{", text)
    depth, j = 1, bm.end()
    while depth and j < len(text):
        depth += {"{": 1, "}": -1}.get(text[j], 0)
        j += 1
    bf = {}
    for name, b in _dict_entries(text[bm.end():j - 1]).items():
        typ = _keyword(b, "type")
        vm = re.search(r"value\s+(.*?);", b, re.S)
        val = None
        if vm is not None and name in patch_sizes:
            val = _field_value(vm.group(1), patch_sizes[name], vector)
        bf[name] = dict(type=typ, value=val, entries=b)
    return internal, bf


def face_geometry(points: np.ndarray, faces: List[np.ndarray]):
    """Face centres and area vectors (primitiveMesh::makeFaceCentresAndAreas)."""
    nf = len(faces)
    Cf = np.zeros((nf, 3))
    Sf = np.zeros((nf, 3))
    for f, lab in enumerate(faces):
        p = points[lab]
        if lab.size == 3:
            Cf[f] = (p[0] + p[1] + p[2]) / 3.0
            Sf[f] = 0.5 * np.cross(p[1] - p[0], p[2] - p[0])
            continue
        fc = p.mean(axis=0)
        q = np.roll(p, -1, axis=0)
        c = p + q + fc
        n = np.cross(q - p, fc - p)
        a = np.linalg.norm(n, axis=1)
        sa = a.sum()
        Cf[f] = (1.0 / 3.0) * (a[:, None] * c).sum(axis=0) / sa if sa > VSMALL else fc
        Sf[f] = 0.5 * n.sum(axis=0)
    return Cf, Sf


def cell_geometry(Cf: np.ndarray, Sf: np.ndarray, owner: np.ndarray, neighbour: np.ndarray, n_cells: int):
    """Cell centres and volumes (primitiveMesh::makeCellCentresAndVols)."""
    nfi = neighbour.size
    cnt = np.bincount(owner, minlength=n_cells) + np.bincount(neighbour, minlength=n_cells)
    cest = np.zeros((n_cells, 3))
    np.add.at(cest, owner, Cf)
    np.add.at(cest, neighbour, Cf[:nfi])
    cest /= cnt[:, None]
    pyr_o = np.einsum("ij,ij->i", Sf, Cf - cest[owner])
    pyr_n = np.einsum("ij,ij->i", Sf[:nfi], cest[neighbour] - Cf[:nfi])
    V = np.zeros(n_cells)
    C = np.zeros((n_cells, 3))
    np.add.at(V, owner, pyr_o)
    np.add.at(V, neighbour, pyr_n)
    np.add.at(C, owner, pyr_o[:, None] * (0.75 * Cf + 0.25 * cest[owner]))
    np.add.at(C, neighbour, pyr_n[:, None] * (0.75 * Cf[:nfi] + 0.25 * cest[neighbour]))
    ok = np.abs(V) > VSMALL
    C[ok] /= V[ok, None]
    C[~ok] = cest[~ok]
    return C, V / 3.0


GREAT = 1.0e15
SMALL = 1.0e-15


def read_controls(path: str) -> dict:
    """system/controlDict entries the solver loop reads, with the reference's defaults
    (readTimeControlsExplicit.H:39-52; controlDict deltaT, endTime, writeControl, writeInterval)."""
    ctl = _read(os.path.join(path, "system", "controlDict"))

    def num(key, default):
        v = _keyword(ctl, key)
        return default if v is None else _number(v)

    def flag(key, default):
        v = _keyword(ctl, key)
        return default if v is None else v in ("yes", "on", "true", "1")

    return dict(deltaT=num("deltaT", 1.0), startTime=num("startTime", 0.0), endTime=num("endTime", GREAT),
                adjustTimeStep=flag("adjustTimeStep", False), maxCo=num("maxCo", 1.0),
                maxDeltaT=num("maxDeltaT", GREAT), convergeTol=num("convergeTol", 1.0e-9),
                convergeCheckSteps=int(num("convergeCheckSteps", 1)),
                writeControl=_keyword(ctl, "writeControl") or "timeStep",
                writeInterval=num("writeInterval", 1.0), timePrecision=int(num("timePrecision", 6)))


def next_delta_t(co: float, dt: float, ctl: dict) -> float:
    """setDeltaTvar.H:34-49: immediate reduction, damped growth, capped at maxDeltaT."""
    f = ctl["maxCo"] / (co + SMALL)
    return min(min(min(f, 1.0 + 0.1 * f), 1.2) * dt, ctl["maxDeltaT"])


def first_delta_t(C: np.ndarray, geo: dict, U: np.ndarray, dim: int, xi_max: float, ctl: dict) -> float:
    """The step the reference's first iteration takes with adjustTimeStep on: the Courant
    number of fvDVM::getCoNum (fvDVM.C:1111-1119) at controlDict's deltaT, with the linearly
    interpolated initial velocity on the faces (fvDVM.C:1074), fed to next_delta_t."""
    own, nei, w = geo["owner"], geo["neighbour"], geo["weights"]
    nfi = nei.size
    if nfi == 0:
        return ctl["deltaT"]
    Uf = w[:nfi, None] * U[own[:nfi]] + (1.0 - w[:nfi, None]) * U[nei]
    dc = 1.0 / np.linalg.norm(C[nei] - C[own[:nfi]], axis=1)
    co = float((dc * (np.linalg.norm(Uf, axis=1) + math.sqrt(dim) * xi_max)).max()) * ctl["deltaT"]
    return next_delta_t(co, ctl["deltaT"], ctl)


def read_case(path: str, dt: float = None) -> cases.Case:
    """A Case from a dugksFoam case directory.  dt defaults to controlDict's deltaT, or with
    adjustTimeStep on, to the step the reference takes first (first_delta_t)."""
    pm = os.path.join(path, "constant", "polyMesh")
    points = read_points(os.path.join(pm, "points"))
    faces = read_faces(os.path.join(pm, "faces"))
    owner = read_labels(os.path.join(pm, "owner"))
    neighbour = read_labels(os.path.join(pm, "neighbour"))
    bnd = read_boundary(os.path.join(pm, "boundary"))
    nfi = neighbour.size
    n_cells = int(max(owner.max(), neighbour.max() if nfi else 0)) + 1
    Cf_all, Sf_all = face_geometry(points, faces)
    C, V = cell_geometry(Cf_all, Sf_all, owner, neighbour, n_cells)

    # 2-D: the directions normal to empty patches carry no solution
    empty_dirs = set()
    for b in bnd:
        if b["type"] == "empty" and b["nFaces"]:
            s = Sf_all[b["startFace"]]
            empty_dirs.add(int(np.argmax(np.abs(s))))
    dim = 3 - len(empty_dirs)
    if dim == 2 and empty_dirs != {2}:
        raise NotImplementedError("2-D cases must be empty in z (the velocity set has xi_z = 0)")
    if dim == 1 and empty_dirs != {1, 2}:
        raise NotImplementedError("1-D cases must be empty in y and z (the velocity set has xi_y = xi_z = 0)")

    sizes = {b["name"]: b["nFaces"] for b in bnd}
    rho_i, rho_b = read_field(os.path.join(path, "0", "rho"), n_cells, sizes, vector=False)
    U_i, U_b = read_field(os.path.join(path, "0", "U"), n_cells, sizes, vector=True)
    T_i, T_b = read_field(os.path.join(path, "0", "T"), n_cells, sizes, vector=False)

    live = [b for b in bnd if b["type"] != "empty" and b["nFaces"] > 0]
    index = {b["name"]: i for i, b in enumerate(live)}
    plist = []
    for b in live:
        s, n = b["startFace"], b["nFaces"]
        sl = slice(s, s + n)
        rtype = rho_b[b["name"]]["type"]
        if b["type"] == "symmetryPlane" or rtype == "symmetryPlane":
            # OpenFOAM constraint type: the mesh patch and rho, U, T are all symmetryPlane
            # (doc/usage.tex:80-81); gSurf follows as symmetryPlane (discreteVelocity.C:257-261)
            types = {b["type"], rtype, U_b[b["name"]]["type"], T_b[b["name"]]["type"]}
            if types != {"symmetryPlane"}:
                raise NotImplementedError(f"patch {b['name']}: a symmetryPlane patch needs the symmetryPlane "
                                          f"type in polyMesh/boundary and in 0/rho, 0/U, 0/T")
            plist.append(cases.Patch(b["name"], cases.PATCH_SYMMETRY_PLANE, owner[sl], Sf_all[sl], Cf_all[sl]))
            continue
        if b["type"] == "cyclic":
            plist.append(cases.Patch(b["name"], cases.PATCH_CYCLIC, owner[sl], Sf_all[sl], Cf_all[sl],
                                     neighbour_patch=index[b["neighbourPatch"]]))
        elif rtype in ("farField", "fixedValue"):
            # farField / mixed DF patches (bcMap, discreteVelocity.C:275-281): incoming Maxwellian
            # of the boundary rho and T; the far field takes U from the boundary cells
            if T_b[b["name"]]["type"] != "fixedValue" or rho_b[b["name"]]["value"] is None:
                raise NotImplementedError(f"patch {b['name']}: {rtype} rho needs a rho value and fixedValue T")
            kw = dict(T=T_b[b["name"]]["value"], rho=rho_b[b["name"]]["value"])
            if rtype == "fixedValue":
                if U_b[b["name"]]["value"] is None:
                    raise NotImplementedError(f"patch {b['name']}: fixedValue rho needs a U value")
                kw["U"] = U_b[b["name"]]["value"]
            kind = cases.PATCH_FARFIELD if rtype == "farField" else cases.PATCH_FIXED
            plist.append(cases.Patch(b["name"], kind, owner[sl], Sf_all[sl], Cf_all[sl], **kw))
        elif rtype in ("pressureIn", "pressureOut"):
            # farField DF patches whose rho (and, at the outlet, T) follow the patch pressure
            # after every step (fvDVM.C:730-806); the pressure is the patch's own keyword
            pk = _keyword(rho_b[b["name"]]["entries"], rtype)
            if pk is None:
                raise NotImplementedError(f"patch {b['name']}: {rtype} without its {rtype} keyword")
            tb = T_b[b["name"]]
            if tb["type"] == "fixedValue":
                Tp = tb["value"]
            elif rtype == "pressureOut" and tb["type"] == "zeroGradient":
                Tp = T_i[owner[sl]]
            else:
                raise NotImplementedError(f"patch {b['name']}: {rtype} needs a fixedValue T")
            kind = cases.PATCH_PRESSURE_IN if rtype == "pressureIn" else cases.PATCH_PRESSURE_OUT
            plist.append(cases.Patch(b["name"], kind, owner[sl], Sf_all[sl], Cf_all[sl], T=Tp,
                                     pressure=_number(pk)))
        elif rtype == "zeroGradient":
            # zeroGradient DF patch (bcMap, discreteVelocity.C:276): face values = the cells' values
            plist.append(cases.Patch(b["name"], cases.PATCH_ZEROGRADIENT, owner[sl], Sf_all[sl], Cf_all[sl]))
        elif rtype == "symmetryMod":
            # DVMsymmetry DF patch (bcMap, discreteVelocity.C:279): mirror plane
            plist.append(cases.Patch(b["name"], cases.PATCH_SYMMETRY, owner[sl], Sf_all[sl], Cf_all[sl]))
        elif rtype == "calculatedMaxwell":
            if U_b[b["name"]]["type"] != "fixedValue" or T_b[b["name"]]["type"] != "fixedValue":
                raise NotImplementedError(f"patch {b['name']}: Maxwell walls need fixedValue U and T")
            plist.append(cases.Patch(b["name"], cases.PATCH_WALL, owner[sl], Sf_all[sl], Cf_all[sl],
                                     U_b[b["name"]]["value"], T_b[b["name"]]["value"]))
        else:
            raise NotImplementedError(f"patch {b['name']}: rho boundary type {rtype} is not supported")
    # the engine numbers boundary faces patch after patch in this order, as fvMesh does
    geo = cases._assemble(os.path.basename(os.path.normpath(path)), dim, C, V, owner[:nfi], neighbour,
                          Sf_all[:nfi], Cf_all[:nfi], plist)

    props = _read(os.path.join(path, "constant", "DVMProperties"))
    sub = _dict_entries(props)
    dv, gas = sub["fvDVMparas"], sub["gasProperties"]
    ndv = int(_number(_keyword(dv, "nDV")))
    xi_max = _number(_keyword(dv, "xiMax"))
    xi = read_labels_float(os.path.join(path, "constant", "Xis"))
    w = read_labels_float(os.path.join(path, "constant", "weights"))
    if xi.size != ndv or w.size != ndv:
        raise ValueError("constant/Xis and constant/weights must hold nDV values")
    R = _number(_keyword(gas, "R"))
    omega = _number(_keyword(gas, "omega"))
    Tref = _number(_keyword(gas, "Tref"))
    mu = _number(_keyword(gas, "muRef"))
    Pr = _number(_keyword(gas, "Pr"))
    kin = _keyword(gas, "KInner")
    K = int(_number(kin)) if kin is not None else 0

    ctl = read_controls(path)
    if dt is None:
        if ctl["adjustTimeStep"]:
            dt = first_delta_t(C, geo, U_i, dim, xi_max, ctl)
        else:
            dt = ctl["deltaT"]
    case = cases._finish(geo, xi, w, xi_max, R, omega, Tref, mu, Pr, K, 1.0, 1.0, dt)
    case.rho0 = np.ascontiguousarray(rho_i, dtype=np.float64)
    case.U0 = np.ascontiguousarray(U_i, dtype=np.float64)
    case.T0 = np.ascontiguousarray(T_i, dtype=np.float64)
    m = re.search(r"DFwriteCellList\s*(\d*)\s*\(([^)]*)\)", props)
    df_cells = [int(x) for x in m.group(2).split()] if m else []
    case.meta.update(source=os.path.abspath(path), controls=ctl, df_cells=df_cells,
                     wall_patches=[q.name for q in plist if q.type == cases.PATCH_WALL],
                     boundary=bnd)
    return case


def read_labels_float(path: str) -> np.ndarray:
    """A scalarList file (`N ( v0 v1 ... )`)."""
    n, body = _list_body(_read(path))
    a = np.array(body.split(), dtype=np.float64)
    if a.size != n:
        raise ValueError(f"{path}: expected {n} values")
    return a
