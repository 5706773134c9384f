"""Air mesh of the bijectivity scaffold for a map whose boundary is one convex loop (every initial map of the
workloads here: the Tutte map of a cut surface puts its boundary on a circle, a height-field disc starts on its
square).

The region is the one the reference triangulates (Scaffold.cpp:27-176): between the map's boundary and a box
around the map's bounding box, grown by margin = (l - 9 l) / (1 - 3) = 4 l and cut into segments of length at most
9 l on each side (l = TriMesh::avgEdgeLen, the rest mesh's mean edge length, Scaffold.cpp:60-90). The reference
meshes it with Triangle ("qYQ": quality, no Steiner points on the boundary). Here the interior is filled with
rings of points offset from the boundary at spacings growing by a factor 1.3 from the boundary's own edge length,
and the point set is triangulated by Delaunay (scipy's Qhull); with a convex boundary and ring points kept
outside the boundary edges' diametral circles, every boundary edge is a Delaunay edge, so dropping the triangles
inside the boundary leaves a conforming air mesh. The box's side points bulge outward by 1e-3 of a segment so the
hull has no collinear points.

After the first build the air mesh is kept: its vertices move with the Newton steps, its rest shape is refreshed
to the current UV after every step, and Lawson flips re-mesh it when its triangles degrade (oc_scaffold_update).
"""
import numpy as np
from scipy.spatial import Delaunay

from optcuts_amd.meshgen import boundary_loop_ccw


def _signed_area(P):
    x, y = P[:, 0], P[:, 1]
    return 0.5 * float((x * np.roll(y, -1) - np.roll(x, -1) * y).sum())


def box_points(UV, avg_edge_len):
    """The reference's bounding box with its segments (Scaffold.cpp:60-90), counter-clockwise from (minX, minY)."""
    seg = avg_edge_len * 3.0 ** 2
    margin = (avg_edge_len - seg) / (1.0 - 3.0)
    mn = UV.min(0) - margin
    mx = UV.max(0) + margin
    nx = int(np.ceil((mx[0] - mn[0]) / seg))
    ny = int(np.ceil((mx[1] - mn[1]) / seg))
    sx, sy = (mx[0] - mn[0]) / nx, (mx[1] - mn[1]) / ny
    pts = [(mn[0] + i * sx, mn[1]) for i in range(nx)]
    pts += [(mx[0], mn[1] + i * sy) for i in range(ny)]
    pts += [(mx[0] - i * sx, mx[1]) for i in range(nx)]
    pts += [(mn[0], mx[1] - i * sy) for i in range(ny)]
    P = np.array(pts)
    # outward bulge of the side points (corners stay): no three hull points collinear
    c = 0.5 * (mn + mx)
    bul = np.zeros(len(P))
    bul[1:nx] = np.sin(np.pi * np.arange(1, nx) / nx)
    bul[nx + 1:nx + ny] = np.sin(np.pi * np.arange(1, ny) / ny)
    bul[nx + ny + 1:2 * nx + ny] = np.sin(np.pi * np.arange(1, nx) / nx)
    bul[2 * nx + ny + 1:] = np.sin(np.pi * np.arange(1, ny) / ny)
    side_n = np.zeros_like(P)
    side_n[:nx] = (0, -1)
    side_n[nx:nx + ny] = (1, 0)
    side_n[nx + ny:2 * nx + ny] = (0, 1)
    side_n[2 * nx + ny:] = (-1, 0)
    P = P + 1e-3 * min(sx, sy) * bul[:, None] * side_n
    return P, mn, mx


def initial_air_mesh(UV, F, avg_edge_len, growth=1.3):
    """(airUV, Fair, bnd): the first len(bnd) air vertices are the mesh boundary vertices bnd (the loop with the
    mesh on its left); the others follow. Raises ValueError when the boundary is not one convex loop."""
    UV = np.asarray(UV, np.float64)
    loop = boundary_loop_ccw(F)
    B = UV[loop]
    nb = len(loop)
    if _signed_area(B) <= 0.0:
        raise ValueError("initial_air_mesh: boundary loop is not counter-clockwise")
    e = np.roll(B, -1, 0) - B
    cr = e[:, 0] * np.roll(e, -1, 0)[:, 1] - e[:, 1] * np.roll(e, -1, 0)[:, 0]
    if (cr < -1e-12 * (np.linalg.norm(e, axis=1) ** 2).max()).any():
        raise ValueError("initial_air_mesh: boundary is not convex")
    box, mn, mx = box_points(UV, avg_edge_len)
    seg = avg_edge_len * 9.0
    s0 = float(np.linalg.norm(e, axis=1).mean())
    # outward vertex normals of the convex loop
    en = np.stack([e[:, 1], -e[:, 0]], 1)
    en /= np.linalg.norm(en, axis=1, keepdims=True)
    vn = en + np.roll(en, 1, 0)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    rings = []
    d, s = 0.0, s0
    while True:
        d += s
        s *= growth
        R = B + d * vn
        # resample the ring at spacing s by arc length
        seglen = np.linalg.norm(np.roll(R, -1, 0) - R, axis=1)
        cum = np.concatenate([[0.0], np.cumsum(seglen)])
        L = cum[-1]
        n = max(3, int(np.round(L / s)))
        t = np.arange(n) * (L / n)
        k = np.searchsorted(cum, t, side="right") - 1
        f = (t - cum[k]) / seglen[k]
        Q = R[k] + f[:, None] * (np.roll(R, -1, 0)[k] - R[k])
        keep = ((Q[:, 0] > mn[0] + 0.5 * s) & (Q[:, 0] < mx[0] - 0.5 * s) &
                (Q[:, 1] > mn[1] + 0.5 * s) & (Q[:, 1] < mx[1] - 0.5 * s))
        if not keep.any() or s > seg:
            break
        rings.append(Q[keep])
    pts = np.concatenate([B] + rings + [box])
    tri = Delaunay(pts).simplices.astype(np.int64)
    if np.unique(tri).size != pts.shape[0]:
        raise ValueError("initial_air_mesh: a point was left out of the triangulation")
    # orient counter-clockwise
    P = pts[tri]
    cr = (P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 1, 1] - P[:, 0, 1]) * (P[:, 2, 0] - P[:, 0, 0])
    tri[cr < 0] = tri[cr < 0][:, [0, 2, 1]]
    # drop the triangles inside the boundary: all three vertices on the boundary and the triangle to the left of
    # the loop (its vertex order follows the loop's)
    inb = (tri < nb).all(1)
    cen = pts[tri].mean(1)
    inside = np.ones(len(tri), bool)
    for i in range(nb):
        a, b = B[i], B[(i + 1) % nb]
        inside &= ((b[0] - a[0]) * (cen[:, 1] - a[1]) - (b[1] - a[1]) * (cen[:, 0] - a[0])) > 0
    Fair = tri[~(inb & inside)]
    # every boundary edge (i+1 -> i seen from the air) must be an air edge
    he = np.concatenate([Fair[:, [0, 1]], Fair[:, [1, 2]], Fair[:, [2, 0]]])
    key = set((he[:, 0] * pts.shape[0] + he[:, 1]).tolist())
    for i in range(nb):
        if ((i + 1) % nb) * pts.shape[0] + i not in key:
            raise ValueError("initial_air_mesh: boundary edge %d is not a Delaunay edge" % i)
    P = pts[Fair]
    ar = 0.5 * ((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) -
                (P[:, 1, 1] - P[:, 0, 1]) * (P[:, 2, 0] - P[:, 0, 0]))
    if (ar <= 1e-14 * s0 * s0).any():
        raise ValueError("initial_air_mesh: degenerate air triangle")
    return pts, Fair.astype(np.int32), loop.astype(np.int32)


def area_threshold(avg_edge_len):
    """airMesh.areaThres_AM (Scaffold.cpp:176): the area of an equilateral triangle of edge avgEdgeLen / 2."""
    eps = avg_edge_len * 0.5
    return np.sqrt(3.0) / 4.0 * eps * eps
