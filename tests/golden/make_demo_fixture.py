"""Stores the reference's demo/cavity case as array fixtures, for machines without the
reference tree.

    python tests/golden/make_demo_fixture.py [/root/reference/demo]

writes
  demo_cavity_dv.npz      demo/cavity's constant/Xis and constant/weights (28-point half-range
                          Gauss-Hermite)
  demo_cavity_case.npz    demo/cavity: the polyMesh (points, faces, owner, neighbour, boundary),
                          the 0/ fields (internal and patch values), DVMProperties, Xis/weights
                          and the time controls
  demo_symmetry_case.npz  demo/testSymmetry/symmetry, the same way (half cavity with a
                          symmetryMod plane)

The files are read with the project's own case reader (dugksfoam_amd.foam); nothing of the
reference is executed.  tests/foam_box.write_fixture_case writes the arrays back out as an
OpenFOAM case directory.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dugksfoam_amd import foam  # noqa: E402


def main(src, name):
    pm = os.path.join(src, "constant", "polyMesh")
    points = foam.read_points(os.path.join(pm, "points"))
    faces = foam.read_faces(os.path.join(pm, "faces"))
    owner = foam.read_labels(os.path.join(pm, "owner"))
    neighbour = foam.read_labels(os.path.join(pm, "neighbour"))
    bnd = foam.read_boundary(os.path.join(pm, "boundary"))
    case = foam.read_case(src)
    sizes = {b["name"]: b["nFaces"] for b in bnd}
    out = dict(points=points, face_sizes=np.array([f.size for f in faces], np.int32),
               face_points=np.concatenate(faces).astype(np.int32), owner=owner.astype(np.int32),
               neighbour=neighbour.astype(np.int32),
               patch_names=np.array([b["name"] for b in bnd]), patch_types=np.array([b["type"] for b in bnd]),
               patch_n=np.array([b["nFaces"] for b in bnd], np.int32),
               patch_start=np.array([b["startFace"] for b in bnd], np.int32))
    for k, vec in (("rho", False), ("U", True), ("T", False)):
        internal, bf = foam.read_field(os.path.join(src, "0", k), case.n_cells, sizes, vector=vec)
        out[f"{k}_internal"] = internal
        out[f"{k}_patch_types"] = np.array([bf.get(b["name"], dict(type="empty"))["type"] for b in bnd])
        for b in bnd:
            v = bf.get(b["name"], {}).get("value")
            if v is not None and b["type"] != "empty":
                out[f"{k}_value_{b['name']}"] = v
    for k in ("xi_max", "R", "omega", "T_ref", "mu_ref", "Pr", "K"):
        out[k] = np.array(getattr(case, k))
    out["nDV"] = np.array(case.xi.size)
    out["xis"], out["weights"] = case.xi, case.xi_weights
    for k, v in case.meta["controls"].items():
        out[f"ctl_{k}"] = np.array(v)
    np.savez_compressed(os.path.join(HERE, f"{name}_case.npz"), **out)
    if name == "demo_cavity":
        np.savez_compressed(os.path.join(HERE, "demo_cavity_dv.npz"), xis=case.xi, weights=case.xi_weights)
    print(f"wrote {name}_case.npz:", case.n_cells, "cells,", len(faces), "faces")


if __name__ == "__main__":
    demo = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/demo"
    main(os.path.join(demo, "cavity"), "demo_cavity")
    main(os.path.join(demo, "testSymmetry", "symmetry"), "demo_symmetry")
