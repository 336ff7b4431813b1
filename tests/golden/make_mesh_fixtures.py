"""Write tests/golden/msh_<name>.npz from the reference's bundled Fluent meshes (MST-CFD/msh/*.msh).

The fixtures are data (node coordinates, face rows, zone table) parsed here with a small reader
written from the reference reader's rules (mesh/MshBlock.cpp:75-264): '(' lines are headers, the
BC type comes from the "(13 (zone first last type nnode)" header, rows are hex "n0 n1 c0 c1",
1-based, c1 = 0 on boundary faces.  Run in a container where /root/reference exists:
    python tests/golden/make_mesh_fixtures.py
"""
import os
import sys

import numpy as np

SRC = "/root/reference/MST-CFD/msh"
OUT = os.path.dirname(os.path.abspath(__file__))
NAMES = ["2d-shockwavepipe-2", "2d-stairW-1", "2d-stair-st-2", "2d-stair-un-5-tri"]


def parse(path):
    lines = open(path, "rb").read().decode().replace("\r", "").split("\n")
    i = 0
    nodes = None
    faces = {}
    zones = []
    ncell = 0
    while i < len(lines):
        ln = lines[i].strip()
        i += 1
        if not ln.startswith("("):
            continue
        head = ln[1:].split(" ", 1)[0]
        if head in ("10", "12", "13"):
            inner = ln[ln.index("(", 1) + 1: ln.index(")")]
            f = [int(t, 16) for t in inner.split()]
            rest = ln[ln.index(")") + 1:].strip()
            if f[0] == 0:
                if head == "10":
                    nodes = np.zeros((f[2], 2))
                elif head == "12":
                    ncell = f[2]
                else:
                    nface = f[2]
                continue
            if head == "12":
                continue
            rows = []
            if rest != "(" and rest != ")(":
                # body opens on the next line "("
                assert lines[i].strip() == "(", (path, ln)
                i += 1
            while not lines[i].strip().startswith(")"):
                rows.append(lines[i].split())
                i += 1
            if head == "10":
                for k, r in enumerate(rows):
                    nodes[f[1] - 1 + k] = [float(r[0]), float(r[1])]
            else:
                zones.append((f[0], f[1] - 1, f[2], f[3]))
                for k, r in enumerate(rows):
                    v = [int(t, 16) for t in r]
                    faces[f[1] - 1 + k] = (v[0] - 1, v[1] - 1, v[2] - 1, (v[3] - 1) if f[3] == 2 else -1, f[3])
    fn = np.array([[faces[k][0], faces[k][1]] for k in range(nface)], np.int32)
    c0 = np.array([faces[k][2] for k in range(nface)], np.int32)
    c1 = np.array([faces[k][3] for k in range(nface)], np.int32)
    ft = np.array([faces[k][4] for k in range(nface)], np.int32)
    n_int = max([z[2] for z in zones if z[3] == 2] + [0])
    return dict(dim=2, xyz=nodes, fnodes=fn, c0=c0, c1=c1, ftype=ft, partner=-np.ones(nface, np.int32),
                n_cells=ncell, n_int=n_int, zones=np.array(zones, np.int32))


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} not present")
    for name in NAMES:
        d = parse(os.path.join(SRC, name + ".msh"))
        np.savez_compressed(os.path.join(OUT, f"msh_{name}.npz"), **d)
        print(name, d["n_cells"], d["xyz"].shape[0], d["c0"].shape[0], d["n_int"])


if __name__ == "__main__":
    main()
