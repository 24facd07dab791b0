"""CPU oracle of the density -> STL isosurface -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; the product never does.

The reference (genpipeline/topology/mesh_export.py:7-67) runs
``skimage.measure.marching_cubes(density, level)`` (scikit-image, unpinned
in requirements.txt:13; not installed in this image) and writes the
triangles with numpy-stl. Neither library can run here, so the reference's
exact triangle lists are parity unpinned. This module restates the method:
marching cubes over the (nx, ny, nz) sample grid, a cube per 2x2x2 block of
samples, a vertex on every cube edge whose end samples straddle the level, at
the linear crossing point p0 + (level - v0) / (v1 - v0) (p1 - p0) from the
edge's lower-coordinate end p0 -- the vertex placement skimage uses (the
fixed end makes the cubes that share an edge compute the same point). The
triangulation of each of the 256 corner sign cases is generated, not
tabulated:

* a corner is inside when its sample is > level;
* on each cube face the edge crossings are joined by segments that keep the
  inside corners on their left when the face is walked counter-clockwise
  seen from outside the cube; a face with two diagonal inside corners keeps
  them apart (two segments, one around each);
* the segments of the six faces close into loops (every crossing has one
  segment in and one out), each loop is fanned into triangles from its
  first crossing, against the loop's direction.

Neighbouring cubes see a shared face with the same corner signs, so they
join its crossings the same way with opposite directions: the surface is
closed and consistently oriented wherever the solid does not touch the
grid's outer faces (there, as in the reference, it stays open). Triangle
normals (right-hand rule) point from inside to outside.

How it is pinned (tests/test_oracle.py): every generated case is closed
(each crossing edge used once in each direction) and reproduces the
single-corner cases; on closed shapes the mesh is watertight with opposite
edge directions, a sampled sphere encloses 4/3 pi r^3 within the
discretisation error with positive signed volume and Euler characteristic
2, a torus has characteristic 0, and every vertex is the exact linear
crossing point of its grid edge.
"""
from __future__ import annotations

import numpy as np

# corner c of a cube at offsets (x, y, z); the C3D8 order of simp_solver.py:255-267
CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                    (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
         (0, 4), (1, 5), (2, 6), (3, 7)]
MAX_TRIS = 5


def _faces():
    """The 6 faces as corner cycles, counter-clockwise seen from outside."""
    out = []
    for axis in range(3):
        for side in (0, 1):
            cs = [c for c in range(8) if CORNERS[c][axis] == side]
            ctr = CORNERS[cs].mean(axis=0)
            normal = np.zeros(3)
            normal[axis] = 1.0 if side else -1.0
            # order by angle around the face normal
            u = np.eye(3)[(axis + 1) % 3]
            v = np.cross(normal, u)
            ang = [np.arctan2((CORNERS[c] - ctr) @ v, (CORNERS[c] - ctr) @ u) for c in cs]
            out.append([cs[i] for i in np.argsort(ang)])
    return out


FACES = _faces()


def _edge_of(a, b):
    for i, (p, q) in enumerate(EDGES):
        if (p, q) == (a, b) or (p, q) == (b, a):
            return i
    raise KeyError((a, b))


def case_triangles(case: int):
    """Triangles (edge-index triples) of corner sign case `case` (bit c set:
    corner c inside)."""
    inside = [(case >> c) & 1 == 1 for c in range(8)]
    nxt = {}
    for f in FACES:
        n = len(f)
        # crossings in counter-clockwise order: (edge, entering inside?)
        cr = []
        for i in range(n):
            a, b = f[i], f[(i + 1) % n]
            if inside[a] != inside[b]:
                cr.append((_edge_of(a, b), inside[b]))
        # each inside arc runs from an entering crossing P to the next leaving
        # crossing Q (counter-clockwise); the segment Q -> P closes it with the
        # inside on its left
        m = len(cr)
        for i in range(m):
            e, entering = cr[i]
            if entering:
                q = cr[(i + 1) % m][0]
                assert q not in nxt
                nxt[q] = e
    tris = []
    seen = set()
    for start in sorted(nxt):
        if start in seen:
            continue
        loop = [start]
        seen.add(start)
        e = nxt[start]
        while e != start:
            loop.append(e)
            seen.add(e)
            e = nxt[e]
        # fanned against the loop direction: normals point inside -> outside
        for i in range(1, len(loop) - 1):
            tris.append((loop[0], loop[i + 1], loop[i]))
    return tris


def tables():
    """(ntri[256] int8, tri[256][MAX_TRIS*3] int8, -1 padded)."""
    ntri = np.zeros(256, np.int8)
    tri = -np.ones((256, 3 * MAX_TRIS), np.int8)
    for case in range(256):
        t = case_triangles(case)
        assert len(t) <= MAX_TRIS
        ntri[case] = len(t)
        for i, (a, b, c) in enumerate(t):
            tri[case, 3 * i:3 * i + 3] = (a, b, c)
    return ntri, tri


def marching_cubes(density: np.ndarray, level: float, scale=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Triangles (T, 3, 3) float64 in physical units (index * scale), cubes
    in C order of their first corner, each cube's triangles in table order
    (at most MAX_TRIS = 5 per cube)."""
    d = np.asarray(density, np.float64)
    nx, ny, nz = d.shape
    ntri, tri = tables()
    out = []
    sc = np.asarray(scale, np.float64)
    for i in range(nx - 1):
        for j in range(ny - 1):
            for k in range(nz - 1):
                v = [d[i + CORNERS[c][0], j + CORNERS[c][1], k + CORNERS[c][2]] for c in range(8)]
                case = sum(1 << c for c in range(8) if v[c] > level)
                for t in range(ntri[case]):
                    pts = []
                    for e in tri[case, 3 * t:3 * t + 3]:
                        a, b = EDGES[e]
                        if CORNERS[a].sum() > CORNERS[b].sum():  # from the lower end
                            a, b = b, a
                        va, vb = v[a], v[b]
                        s = (level - va) / (vb - va)
                        pa = np.array([i, j, k], np.float64) + CORNERS[a]
                        pb = np.array([i, j, k], np.float64) + CORNERS[b]
                        pts.append((pa + s * (pb - pa)) * sc)
                    out.append(pts)
    return np.array(out, np.float64).reshape(-1, 3, 3)


def effective_level(density: np.ndarray, threshold: float = 0.5):
    """density_to_stl's level (mesh_export.py:38-49): None for an empty
    design (max < 0.01); a threshold outside [min, max] moves 10 % into the
    range; then at least 0.1."""
    d = np.asarray(density)
    if d.max() < 0.01:
        return None
    lo, hi = float(d.min()), float(d.max())
    if threshold < lo:
        threshold = lo + 0.1 * (hi - lo)
    if threshold > hi:
        threshold = hi - 0.1 * (hi - lo)
    return max(threshold, 0.1)


def stl_records(tris: np.ndarray) -> np.ndarray:
    """(T, 12) float32 STL records: unit normal cross(v1 - v0, v2 - v0) /
    |...| (0 for a degenerate triangle), then the three vertices; computed in
    float64 from the float64 vertices, rounded once."""
    t = np.asarray(tris, np.float64)
    a = t[:, 1] - t[:, 0]
    b = t[:, 2] - t[:, 0]
    n = np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                  a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                  a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)
    ln = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
    with np.errstate(invalid="ignore", divide="ignore"):
        n = np.where(ln[:, None] > 0, n / np.where(ln > 0, ln, 1.0)[:, None], 0.0)
    return np.concatenate([n, t.reshape(-1, 9)], axis=1).astype(np.float32)
