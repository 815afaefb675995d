"""Cell and face geometry of a structured block from its grid vertices: the
host half of SFluidBlock.compute_primary_cell_geometric_data
(eilmer/sfluidblock.d:650-842) for blocks whose cells are not uniform boxes.

Every formula keeps the reference's operation order, evaluated elementwise
over whole arrays, so each value is the same double the reference's scalar
code computes:
  hexahedron volume (six pyramids on the faces about the vertex mean, each as
      four tetrahedra about its base mid-point) and the i, j, k widths between
      opposite face mid-points                       geom/elements/properties.d:568-650, 293-316, 390-419
  quadrilateral face area and unit normal, t1 along the p0->p1 direction,
      t2 = n x t1                                    geom/elements/properties.d:97-139
  2D cell area and widths                            geom/elements/properties.d:177-237
  2D face length and normal (to the right of A->B)   eilmer/fvinterface.d:265-297
with the cell vertex order of sfluidblock.d:486-645 (p0..p7 = (i,j,k),
(i+1,j,k), (i+1,j+1,k), (i,j+1,k), then the same at k+1), i-faces through
(i,j,k) (i,j+1,k) (i,j+1,k+1) (i,j,k+1), j-faces through (i,j,k) (i,j,k+1)
(i+1,j,k+1) (i+1,j,k) and k-faces through (i,j,k) (i+1,j,k) (i+1,j+1,k)
(i,j+1,k).

The result is the `geom` array of e4hip_block_create_geometry: cells
[4][ncells] (volume, iLength, jLength, kLength), then for each direction d of
the block faces [10][nfaces(d)] (area, n.x, n.y, n.z, t1.x, t1.y, t1.z, t2.x,
t2.y, t2.z); cells i fastest, faces of direction d i fastest over
(ni + [d==0]) x (nj + [d==1]) x (nk + [d==2]).
"""
import numpy as np


def _normalize(v):
    """vector3.d:386-407: unit vector, components below 1e-30 flushed to 0."""
    x, y, z = v
    mag = np.sqrt(x * x + y * y + z * z)
    pos = mag > 0.0
    safe = np.where(pos, mag, 1.0)
    out = [np.where(pos, c / safe, 0.0) for c in (x, y, z)]
    return [np.where(np.abs(c) < 1.0e-30, 0.0, c) for c in out]


def _cross(a, b):
    """vector3.d:604-609."""
    return [a[1] * b[2] - b[1] * a[2], b[0] * a[2] - a[0] * b[2], a[0] * b[1] - b[0] * a[1]]


def quad_properties(p0, p1, p2, p3):
    """properties.d:97-139 over arrays of points [..., 3]: (area, n, t1, t2),
    vectors as lists of three arrays."""
    p01 = [p1[..., q] - p0[..., q] + p2[..., q] - p3[..., q] for q in range(3)]
    p03 = [p3[..., q] - p0[..., q] + p2[..., q] - p1[..., q] for q in range(3)]
    va = [p01[1] * p03[2] - p01[2] * p03[1], p01[2] * p03[0] - p01[0] * p03[2], p01[0] * p03[1] - p01[1] * p03[0]]
    va = [c * 0.25 for c in va]
    area = np.sqrt(va[0] * va[0] + va[1] * va[1] + va[2] * va[2])
    ok = area > 1.0e-20
    safe = np.where(ok, area, 1.0)
    n = [c / safe for c in va]
    t1 = _normalize(p01)
    t2 = _normalize(_cross(n, t1))
    one, zero = np.ones_like(area), np.zeros_like(area)
    defaults = ([one, zero, zero], [zero, one, zero], [zero, zero, one])
    n, t1, t2 = ([np.where(ok, c, d) for c, d in zip(vec, dflt)] for vec, dflt in zip((n, t1, t2), defaults))
    return area, n, t1, t2


def _tet_volume(p0, p1, p2, p3):
    """properties.d:293-316 (volume only)."""
    dx01, dy01, dz01 = p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]
    dx02, dy02, dz02 = p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]
    cx = dy01 * dz02 - dz01 * dy02
    cy = dz01 * dx02 - dx01 * dz02
    cz = dx01 * dy02 - dy01 * dx02
    dx03, dy03, dz03 = p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]
    return (dx03 * cx + dy03 * cy + dz03 * cz) / 6.0


def _pyramid_volume(p0, p1, p2, p3, p4):
    """properties.d:390-419: four tetrahedra about the base mid-point."""
    pmB = [0.25 * (p0[q] + p1[q] + p2[q] + p3[q]) for q in range(3)]
    base = (p0, p1, p2, p3, p0)
    volume = 0.0
    for t in range(4):
        volume = volume + _tet_volume(base[t], base[t + 1], pmB, p4)
    return volume


def hex_cell_properties(p):
    """properties.d:568-650 over arrays: p is a list of the eight vertices,
    each a list of three coordinate arrays.  Returns (volume, iLen, jLen, kLen)."""
    c = [0.125 * (p[0][q] + p[1][q] + p[2][q] + p[3][q] + p[4][q] + p[5][q] + p[6][q] + p[7][q]) for q in range(3)]

    def mid(a, b, cc, d):
        return [0.25 * (p[a][q] + p[b][q] + p[cc][q] + p[d][q]) for q in range(3)]

    pmN, pmE, pmS = mid(3, 2, 6, 7), mid(1, 2, 6, 5), mid(0, 1, 5, 4)
    pmW, pmT, pmB = mid(0, 3, 7, 4), mid(4, 5, 6, 7), mid(0, 1, 2, 3)

    def dist(a, b):
        dx, dy, dz = a[0] - b[0], a[1] - b[1], a[2] - b[2]
        return np.sqrt(dx * dx + dy * dy + dz * dz)

    iLen, jLen, kLen = dist(pmE, pmW), dist(pmN, pmS), dist(pmT, pmB)
    pyr = ((6, 7, 3, 2), (5, 6, 2, 1), (4, 5, 1, 0), (7, 4, 0, 3), (7, 6, 5, 4), (0, 1, 2, 3))
    volume = 0.0
    for q in pyr:
        volume = volume + _pyramid_volume(p[q[0]], p[q[1]], p[q[2]], p[q[3]], c)
    volume = np.where(((volume < 0.0) & (np.abs(volume) < 1.0e-12)) | ((volume >= 0.0) & (volume < 1.0e-20)),
                      0.0, volume)
    return volume, iLen, jLen, kLen


def xyplane_quad_cell_properties(p0, p1, p2, p3):
    """properties.d:177-237: (area, iLen, jLen) of the planar quadrilateral
    cells p0 p1 p2 p3 (points as lists of coordinate arrays)."""
    xA, yA, xB, yB, xC, yC, xD, yD = p1[0], p1[1], p2[0], p2[1], p3[0], p3[1], p0[0], p0[1]
    area = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) + (xD + xC) * (yD - yC) + (xA + xD) * (yA - yD))
    xN, yN = 0.5 * (xC + xB), 0.5 * (yC + yB)
    xS, yS = 0.5 * (xD + xA), 0.5 * (yD + yA)
    xE, yE = 0.5 * (xA + xB), 0.5 * (yA + yB)
    xW, yW = 0.5 * (xD + xC), 0.5 * (yD + yC)
    dx, dy = xN - xS, yN - yS
    jLen = np.sqrt(dx * dx + dy * dy)
    dx, dy = xE - xW, yE - yW
    iLen = np.sqrt(dx * dx + dy * dy)
    return area, iLen, jLen


def face_2d_properties(A, B):
    """fvinterface.d:265-297 (planar): (area, n, t1, t2) of the faces from A
    to B, normal to the right."""
    dx, dy = B[0] - A[0], B[1] - A[1]
    LAB = np.sqrt(dx * dx + dy * dy)
    ok = LAB > 1.0e-12
    safe = np.where(ok, LAB, 1.0)
    zero, one = np.zeros_like(LAB), np.ones_like(LAB)
    n = [(B[1] - A[1]) / safe, -(B[0] - A[0]) / safe, zero]
    t2 = [zero, zero, one]
    t1 = _cross(n, t2)
    defaults = ([one, zero, zero], [zero, one, zero], [zero, zero, one])
    n, t1, t2 = ([np.where(ok, c, d) for c, d in zip(vec, dflt)] for vec, dflt in zip((n, t1, t2), defaults))
    return np.where(ok, LAB, 0.0), n, t1, t2


def _pack_faces(area, n, t1, t2):
    return np.stack([area] + list(n) + list(t1) + list(t2)).reshape(10, -1)


def block_geometry(grid):
    """The geometry array of a block from its vertices: grid
    [nk+1][nj+1][ni+1][3] (3D) or [nj+1][ni+1][2 or 3] (2D).  Returns
    (dims, ni, nj, nk, geom) with geom laid out as the module docstring says.
    Raises ValueError when a cell has no positive volume (area in 2D)."""
    g = np.asarray(grid, dtype=np.float64)
    if g.ndim == 4:
        nk, nj, ni = g.shape[0] - 1, g.shape[1] - 1, g.shape[2] - 1
        X = [g[..., q] for q in range(3)]

        def vtx(oi, oj, ok, ei=0, ej=0, ek=0):
            # vertices (i + oi, j + oj, k + ok) over the index ranges of the
            # cells (ei/ej/ek = 1 extends a range by one for faces)
            return [X[q][ok:ok + nk + ek, oj:oj + nj + ej, oi:oi + ni + ei] for q in range(3)]

        p = [vtx(0, 0, 0), vtx(1, 0, 0), vtx(1, 1, 0), vtx(0, 1, 0),
             vtx(0, 0, 1), vtx(1, 0, 1), vtx(1, 1, 1), vtx(0, 1, 1)]
        vol, iL, jL, kL = hex_cell_properties(p)
        if not np.all(vol > 0.0):
            raise ValueError("a cell of the grid has no positive volume")
        cells = np.stack([vol, iL, jL, kL]).reshape(4, -1)

        def pts(*vs):
            return [np.stack(v, axis=-1) for v in vs]

        fi = quad_properties(*pts(vtx(0, 0, 0, ei=1), vtx(0, 1, 0, ei=1), vtx(0, 1, 1, ei=1), vtx(0, 0, 1, ei=1)))
        fj = quad_properties(*pts(vtx(0, 0, 0, ej=1), vtx(0, 0, 1, ej=1), vtx(1, 0, 1, ej=1), vtx(1, 0, 0, ej=1)))
        fk = quad_properties(*pts(vtx(0, 0, 0, ek=1), vtx(1, 0, 0, ek=1), vtx(1, 1, 0, ek=1), vtx(0, 1, 0, ek=1)))
        parts = [cells] + [_pack_faces(*f) for f in (fi, fj, fk)]
        return 3, ni, nj, nk, np.concatenate([a.ravel() for a in parts])
    if g.ndim != 3:
        raise ValueError("grid must be [nk+1][nj+1][ni+1][3] or [nj+1][ni+1][2|3]")
    nj, ni = g.shape[0] - 1, g.shape[1] - 1
    X = [g[..., 0], g[..., 1]]

    def vtx2(oi, oj, ei=0, ej=0):
        return [X[q][oj:oj + nj + ej, oi:oi + ni + ei] for q in range(2)]

    area, iL, jL = xyplane_quad_cell_properties(vtx2(0, 0), vtx2(1, 0), vtx2(1, 1), vtx2(0, 1))
    if not np.all(area > 0.0):
        raise ValueError("a cell of the grid has no positive area")
    cells = np.stack([area, iL, jL, np.zeros_like(area)]).reshape(4, -1)
    fi = face_2d_properties(vtx2(0, 0, ei=1), vtx2(0, 1, ei=1))
    fj = face_2d_properties(vtx2(1, 0, ej=1), vtx2(0, 0, ej=1))
    parts = [cells] + [_pack_faces(*f) for f in (fi, fj)]
    return 2, ni, nj, 1, np.concatenate([a.ravel() for a in parts])


def split_geometry(dims, ni, nj, nk, geom):
    """(cells [4][ncells], [faces [10][nfaces(d)] for each direction]) views of a geometry array."""
    nk = nk if dims == 3 else 1
    nc = ni * nj * nk
    cells = geom[:4 * nc].reshape(4, nc)
    faces, off = [], 4 * nc
    for d in range(dims):
        nf = (ni + (d == 0)) * (nj + (d == 1)) * (nk + (d == 2))
        faces.append(geom[off:off + 10 * nf].reshape(10, nf))
        off += 10 * nf
    return cells, faces
