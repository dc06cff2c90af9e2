"""Closed iso-surface of a density field and binary STL output: the design
export of TOP3D_XL.m:206-222 and IO_ExportDesignInTriSurface_stl (572-577).

The reference takes the xPhys field on the padded (resY, resX, resZ) grid,
flips it along y (209), extracts the 0.5 level set with MATLAB's
`isosurface` plus the boundary `isocaps` (212-213) and writes both as one
triangulation with `stlwrite`.  Here the level set is the exact 0.5 level set
of the piecewise-linear interpolant of the samples on the Freudenthal (Kuhn)
split of every grid cube into 6 tetrahedra sharing the cube's main diagonal,
a conforming split, so neighbouring cubes agree on their shared faces and the
surface has no cracks.  The caps come from one padding layer of a very
negative value around the field: the surface then closes on the outer grid
planes exactly where the field is above the level, as `isocaps` does.
Triangles face outwards (away from the region above the level).

The triangulation differs from MATLAB's marching-cubes tables (which
triangulate the same level set differently inside each cube); what is fixed
is the level set of the interpolant: linear fields are reproduced exactly,
and the enclosed volume and surface area follow from it.
"""
from __future__ import annotations

import numpy as np

# Freudenthal split of the unit cube: corner c = (dy, dx, dz) bits, 6 tets
# along the c000-c111 diagonal (one per permutation of the three axes).
_CORNERS = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)])
_TETS = []
for axes in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    path, cur = [0], 0
    for a in axes:
        cur |= 1 << a
        path.append(cur)
    _TETS.append(path)
_TETS = np.array(_TETS)  # (6, 4) corner ids

# Case table: for each inside-mask of the 4 tet vertices, the triangles as
# lists of 3 tet edges (vertex pairs).
_CASES = {}
for code in range(16):
    ins = [i for i in range(4) if (code >> i) & 1]
    out = [i for i in range(4) if not (code >> i) & 1]
    tris = []
    if len(ins) == 1 or len(out) == 1:
        a = ins[0] if len(ins) == 1 else out[0]
        others = [i for i in range(4) if i != a]
        tris.append([(a, others[0]), (a, others[1]), (a, others[2])])
    elif len(ins) == 2:
        a, b = ins
        c, d = out
        tris.append([(a, c), (a, d), (b, d)])
        tris.append([(a, c), (b, d), (b, c)])
    _CASES[code] = tris


def isosurface_with_caps(V, iso=0.5, pad_value=-1e30):
    """Triangles of the closed surface around {V >= iso}.

    V is sampled at integer grid points; a vertex at V[i, j, k] (0-based) has
    the coordinates (x, y, z) = (j + 1, i + 1, k + 1), MATLAB's isosurface
    convention.  Returns an (m, 3, 3) float64 array of triangles (vertex
    rows), oriented outwards; triangles of zero area are dropped.
    """
    V = np.asarray(V, dtype=np.float64)
    W = np.full(tuple(s + 2 for s in V.shape), pad_value)
    W[1:-1, 1:-1, 1:-1] = V
    ny, nx, nz = W.shape
    inside = W >= iso
    out = []
    for k in range(nz - 1):
        # cubes of this slab with a sign change (min below, max at or above)
        a = inside[:, :, k:k + 2]
        anyin = np.zeros((ny - 1, nx - 1), bool)
        allin = np.ones((ny - 1, nx - 1), bool)
        for c in range(8):
            dy, dx, dz = _CORNERS[c]
            s = a[dy:dy + ny - 1, dx:dx + nx - 1, dz]
            anyin |= s
            allin &= s
        iy, ix = np.nonzero(anyin & ~allin)
        if iy.size == 0:
            continue
        # corner values and positions of the selected cubes
        cv = np.empty((iy.size, 8))
        cp = np.empty((iy.size, 8, 3))
        for c in range(8):
            dy, dx, dz = _CORNERS[c]
            cv[:, c] = W[iy + dy, ix + dx, k + dz]
            # padded index p -> field index p - 1 -> coordinate p (1-based)
            cp[:, c, 0] = ix + dx
            cp[:, c, 1] = iy + dy
            cp[:, c, 2] = k + dz
        for t in _TETS:
            tv = cv[:, t]                      # (n, 4)
            tp = cp[:, t]                      # (n, 4, 3)
            code = ((tv >= iso) * (1 << np.arange(4))).sum(axis=1)
            for cd in range(1, 15):
                sel = np.nonzero(code == cd)[0]
                if sel.size == 0:
                    continue
                v, p = tv[sel], tp[sel]
                inner = np.array([(cd >> i) & 1 for i in range(4)], bool)
                # outward direction: from the inside vertices to the outside ones
                dirv = p[:, ~inner].mean(axis=1) - p[:, inner].mean(axis=1)
                for tri in _CASES[cd]:
                    pts = []
                    for (i, j) in tri:
                        tt = (iso - v[:, i]) / (v[:, j] - v[:, i])
                        pts.append(p[:, i] + tt[:, None] * (p[:, j] - p[:, i]))
                    T = np.stack(pts, axis=1)  # (n, 3, 3)
                    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
                    flip = (nrm * dirv).sum(axis=1) < 0
                    T[flip] = T[flip][:, ::-1]
                    keep = np.linalg.norm(nrm, axis=1) > 0
                    out.append(T[keep])
    if not out:
        return np.zeros((0, 3, 3))
    return np.concatenate(out)


def write_stl(path, tris):
    """Binary STL of an (m, 3, 3) triangle array (unit normals from the
    vertex order)."""
    tris = np.asarray(tris, dtype=np.float64)
    nrm = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.divide(nrm, ln, out=np.zeros_like(nrm), where=ln > 0)
    rec = np.zeros(len(tris), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    rec["n"] = nrm
    rec["v"] = tris
    with open(path, "wb") as f:
        f.write(b"t3x design iso-surface".ljust(80, b" "))
        f.write(np.uint32(len(tris)).tobytes())
        f.write(rec.tobytes())


def read_stl(path):
    with open(path, "rb") as f:
        raw = f.read()
    n = int(np.frombuffer(raw, "<u4", 1, 80)[0])
    rec = np.frombuffer(raw, dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")], count=n, offset=84)
    return rec["v"].astype(np.float64), rec["n"].astype(np.float64)


def write_design_stl(path, xPhys, cell_map, res, iso=0.5):
    """DesignVolume.stl (206-222): xPhys scattered into the padded grid,
    flipped along y, closed 0.5 iso-surface.  Returns the triangles."""
    from .results import padded_volume
    V = padded_volume(xPhys, cell_map, res)[::-1, :, :]  # flip(isovals, 1), 209
    tris = isosurface_with_caps(V, iso)
    write_stl(path, tris)
    return tris


def enclosed_volume(tris):
    """Signed volume enclosed by a closed, outward-oriented triangulation."""
    tris = np.asarray(tris)
    return float(np.einsum("ij,ij->i", tris[:, 0], np.cross(tris[:, 1], tris[:, 2])).sum() / 6.0)


def surface_area(tris):
    tris = np.asarray(tris)
    return float(0.5 * np.linalg.norm(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]), axis=1).sum())
