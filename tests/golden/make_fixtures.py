"""Extract the reference's committed result files into small numpy fixtures.

Run in the development container (the reference checkout is not present on
the GPU machine):  python tests/golden/make_fixtures.py [REFERENCE_ROOT]

Inputs (reference data/, ASCII outputs of the reference's own experiments):
  solverTest.vtk         Main.cpp:14-38   elastic 4x4 rectangle, CPU path
  holePlatePlastPow.vtk  Main.cpp:252-296 48x32 hole plate, power law, S=170
  ring.vtk               Main.cpp:75-80   genRing(0.1, 0.2, 40, 5) mesh with
                         per-element space / aspect ratio / skew angle
  diagramLin.txt / diagramPow.txt  Main.cpp:929-994  uniaxial sigma(eps)
  rect1.vtk, hole1.vtk, ring1.vtk  Main.cpp:1173-1190  order-2 (8-node) meshes of
                         genRectangle / genRectWithHole / genRing
  renumMesh.vtk          the 48x32 order-2 hole plate mesh (5120 8-node quads)
  aitken_*.txt           Main.cpp:105-178 energyConvTest: deformEnergy per mesh
                         of genRectWithHole(-3, 3, -2, 2, 0.5, N1, N2, order)
Outputs: tests/golden/*.npz (data only: node coordinates, connectivity,
per-element fields and nodal displacements exactly as printed) and
tests/golden/vtk_layout.json (the section header lines of the three VTK
files, in order).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_vtk(path):
    with open(path) as f:
        tok = f.read().split("\n")
    out = {"fields": {}}
    i = 0
    while i < len(tok):
        line = tok[i].strip()
        if line.startswith("POINTS"):
            n = int(line.split()[1])
            pts = np.array([list(map(float, tok[i + 1 + k].split()))[:2] for k in range(n)])
            out["points"] = pts
            i += n + 1
            continue
        if line.startswith("POLYGONS"):
            n = int(line.split()[1])
            polys = np.array([list(map(int, tok[i + 1 + k].split()))[1:] for k in range(n)], dtype=np.int32)
            out["polys"] = polys
            i += n + 1
            continue
        parts = line.split()
        if len(parts) == 4 and parts[3] == "float" and parts[0] not in ("POINTS",):
            name, ncomp, n = parts[0], int(parts[1]), int(parts[2])
            vals = np.array(list(map(float, tok[i + 1].split())))
            assert vals.size == ncomp * n, (name, vals.size)
            out["fields"][name] = vals.reshape(n, ncomp) if ncomp > 1 else vals
            i += 2
            continue
        i += 1
    return out


def main(ref_root):
    data = os.path.join(ref_root, "data")
    for name in ("solverTest", "holePlatePlastPow", "ring"):
        v = read_vtk(os.path.join(data, name + ".vtk"))
        arrays = {"points": v["points"], "polys": v["polys"]}
        for k, a in v["fields"].items():
            arrays["f_" + k] = a
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, {k: a.shape for k, a in arrays.items()})
    # the section layout (every non-numeric line, in order) of the reference's
    # VTK writers: PlasticitySolver::saveAsVtk after solveElast / solve, and
    # Mesh::saveAsVtk after printAnalysis
    layout = {}
    for name in ("solverTest", "holePlatePlastPow", "ring"):
        with open(os.path.join(data, name + ".vtk")) as f:
            layout[name] = [ln.rstrip("\n") for ln in f if ln[:1].isalpha() or ln.startswith("#")]
    with open(os.path.join(HERE, "vtk_layout.json"), "w") as f:
        json.dump(layout, f, indent=0)
    print("vtk_layout", {k: len(v) for k, v in layout.items()})
    for name in ("rect1", "hole1", "ring1", "renumMesh"):
        v = read_vtk(os.path.join(data, name + ".vtk"))
        np.savez_compressed(os.path.join(HERE, "mesh_" + name + ".npz"), points=v["points"], polys=v["polys"])
        print(name, v["points"].shape, v["polys"].shape)
    tables = {}
    for fn in sorted(os.listdir(data)):
        if fn.startswith("aitken") and fn.endswith(".txt"):
            with open(os.path.join(data, fn)) as f:
                rows = [ln.split() for ln in f.read().splitlines()[1:] if ln.strip()]
            tables[fn[:-4]] = [[int(r[0]), r[1]] for r in rows]   # elements, energy as printed
    with open(os.path.join(HERE, "aitken_energy.json"), "w") as f:
        json.dump(tables, f, indent=0)
    print("aitken tables", sorted(tables))
    lin = np.loadtxt(os.path.join(data, "diagramLin.txt"))
    pw = np.loadtxt(os.path.join(data, "diagramPow.txt"))
    np.savez_compressed(os.path.join(HERE, "diagram.npz"), lin=lin, pow=pw)
    print("diagram", lin.shape, pw.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
