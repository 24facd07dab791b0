"""CalculiX deck of a voxel field -- TEST INFRASTRUCTURE ONLY (the checker
of genpipeline_amd.fem.voxel_hex; only tests/ import it).

Restates VoxelHexMesher.voxels_to_inp (genpipeline/fem/voxel_fem.py:76-226)
with its own loops: solid voxels (> 0.5, compared in float32 as the SIMD
scanner of mesher_simd.cpp:11-52 compares) in C order, one C3D8 element
each; grid corners numbered on first touch through a dictionary, corners
in C3D8 order (:120-146); face node sets in numbering order, 16 numbers a
line; the deck's cards in the reference's order (:181-225).

Pinned by tests/test_voxel_hex_cpu.py: byte-equal to decks written by the
reference itself (tests/golden/make_golden_inp.py), the reference's node
count KATs (a 2x2x2 solid cube has 27 nodes, one voxel 8;
tests/test_voxel_fem.py:107-135), and its error messages.
"""
from __future__ import annotations

import numpy as np

C3D8 = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def mesh(voxels):
    """(elements as lists of 8 node numbers, node corners in numbering
    order, element voxels) by first-touch numbering."""
    v = np.asarray(voxels, dtype=np.float32)
    D, H, W = v.shape
    node_of, corners, elems, evox = {}, [], [], []
    for i in range(D):
        for j in range(H):
            for k in range(W):
                if not v[i, j, k] > np.float32(0.5):
                    continue
                ids = []
                for cx, cy, cz in C3D8:
                    key = (i + cx, j + cy, k + cz)
                    if key not in node_of:
                        node_of[key] = len(corners) + 1
                        corners.append(key)
                    ids.append(node_of[key])
                elems.append(ids)
                evox.append((i, j, k))
    return elems, corners, evox


def deck(voxels, bbox=None, fixed_face="x_min", load_face="x_max", force_n=1000.0,
         E_mpa=210000.0, poisson=0.30) -> str:
    """The .inp text; ValueError as the reference raises it."""
    v = np.asarray(voxels)
    D, H, W = v.shape
    if not (v > 0.5).any():
        raise ValueError("All voxels are void — cannot mesh empty geometry.")
    if bbox is None:
        dx = dy = dz = 1.0
        x0 = y0 = z0 = 0.0
    else:
        x0, x1 = bbox["x"]
        y0, y1 = bbox["y"]
        z0, z1 = bbox["z"]
        dx, dy, dz = (x1 - x0) / D, (y1 - y0) / H, (z1 - z0) / W
    elems, corners, _ = mesh(v)
    if not elems:
        raise ValueError("No solid voxels found to mesh.")
    lim = {"x_min": (0, 0), "x_max": (0, D), "y_min": (1, 0), "y_max": (1, H),
           "z_min": (2, 0), "z_max": (2, W)}

    def on(face):
        if face not in lim:
            return []
        ax, val = lim[face]
        return [n + 1 for n, c in enumerate(corners) if c[ax] == val]

    fixed, load = on(fixed_face), on(load_face)
    if not fixed:
        raise ValueError(f"No nodes on fixed face '{fixed_face}'.")
    if not load:
        raise ValueError(f"No nodes on load face '{load_face}'.")
    fpn = force_n / len(load)
    out = ["** CalculiX C3D8 hex mesh — generated by voxel_fem.py", "*NODE, NSET=NALL"]
    for n, (cx, cy, cz) in enumerate(corners, 1):
        out.append("%d,%.6f,%.6f,%.6f" % (n, x0 + cx * dx, y0 + cy * dy, z0 + cz * dz))
    out.append("*ELEMENT, TYPE=C3D8, ELSET=EALL")
    for e, ids in enumerate(elems, 1):
        out.append(str(e) + "," + ",".join(map(str, ids)))
    out += ["*MATERIAL, NAME=STEEL", "*ELASTIC", f"{E_mpa},{poisson}",
            "*SOLID SECTION, ELSET=EALL, MATERIAL=STEEL", "**"]
    for title, ids in (("NFIX", fixed), ("NLOAD", load)):
        out.append(f"*NSET, NSET={title}")
        for a in range(0, len(ids), 16):
            out.append(",".join(map(str, ids[a:a + 16])))
    out += ["**", "*STEP", "*STATIC", "*BOUNDARY", "NFIX,1,3,0.0", "*CLOAD"]
    out += ["%d,3,%.6f" % (n, -fpn) for n in load]
    out += ["*NODE FILE", "U", "*EL FILE", "S", "*END STEP"]
    return "\n".join(out) + "\n"
