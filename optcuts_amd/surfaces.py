"""Closed-surface workloads and their cut to a disc.

Perturbed icosphere: the class-I geodesic subdivision of the icosahedron with frequency nu (20 nu^2 faces,
10 nu^2 + 2 vertices), every vertex on the sphere, radius r(d) = 1 + amp * sum_k exp(-theta(d, d_k)^2 / (2 w_k^2))
over n_lobes directions d_k uniform on the sphere (normalised standard normals) with widths w_k ~ U(lo, hi)
radians; draws from numpy.random.default_rng(seed): the directions, then the widths. nu = 71 gives 100,820
faces, nu = 224 gives 1,003,520.

Fused tori (genus 3): the union of three tori (major radius R = 1, minor r = 0.35, axis z, centres at
x = -1.8, 0, 1.8) as the zero set of the minimum of their signed distances. The surface is extracted by surface
nets on a regular grid (one vertex per grid cell the surface crosses, placed on the surface near the mean of the
cell's edge crossings; one quad per crossing grid edge, split along its shorter diagonal), then relaxed
tangentially (each vertex moves toward the centroid of its neighbours within its tangent plane) and projected
back onto the zero set, which makes the triangles close to isotropic. Finally every vertex is jittered by
N(0, (jitter * mean edge)^2) per coordinate (default_rng(seed)).

Cut to a disc (what main.cpp:1396-1484 does before the Tutte map, with cuts of this module's own):
- genus 0: TriMesh::farthestPointCut(F(0,0)) (TriMesh.cpp:1250-1278): on the vertex graph weighted by rest edge
  lengths, p = the vertex farthest from F[0,0], the path = the shortest path from p to the vertex farthest from p;
- higher genus: a tree-cotree cut graph (the edges not crossed by a spanning tree of the dual graph, built
  breadth-first from triangle 0, with dangling branches pruned) in place of igl::cut_to_disk.
The mesh is cut along the chosen edges by splitting every vertex into one copy per fan of its triangles between
cut edges (triangle corners joined across uncut edges). The cohesive edges follow TriMesh's constructor from a
cut UV layout (TriMesh.cpp:115-182): per cut edge (a, b) one row [a', b', a'', b''] of the copies of a and b on
its two sides, rest lengths from the uncut mesh.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg


# ------------------------------------------------------------------------------------------------------------
# perturbed icosphere

def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return V, F


def geodesic_sphere(nu):
    """Class-I geodesic sphere: (V (10 nu^2 + 2, 3) on the unit sphere, F (20 nu^2, 3) outward CCW)."""
    V0, F0 = _icosahedron()
    # lattice points (i, j) of a face (a, b, c): a + (b - a) i/nu + (c - a) j/nu, i + j <= nu
    ii, jj = np.meshgrid(np.arange(nu + 1), np.arange(nu + 1), indexing="ij")
    keep = ii + jj <= nu
    ii, jj = ii[keep], jj[keep]
    lat = -np.ones((nu + 1, nu + 1), np.int64)
    lat[ii, jj] = np.arange(ii.size)
    # canonical key of each (face, lattice point): icosahedron vertex, edge point or face point
    keys, pos = {}, []

    def key_of(f, i, j):
        a, b, c = (int(x) for x in F0[f])
        k = nu - i - j
        if i == nu:
            return ("v", b)
        if j == nu:
            return ("v", c)
        if k == nu:
            return ("v", a)
        if j == 0:
            return ("e", a, b, i) if a < b else ("e", b, a, nu - i)
        if i == 0:
            return ("e", a, c, j) if a < c else ("e", c, a, nu - j)
        if k == 0:
            return ("e", b, c, j) if b < c else ("e", c, b, nu - j)
        return ("f", f, i, j)

    def point(kk):
        if kk[0] == "v":
            return V0[kk[1]]
        if kk[0] == "e":
            p, q, t = V0[kk[1]], V0[kk[2]], kk[3] / nu
            return p + (q - p) * t
        a, b, c = V0[F0[kk[1]]]
        return a + (b - a) * (kk[2] / nu) + (c - a) * (kk[3] / nu)

    faces = []
    for f in range(20):
        ids = np.empty(ii.size, np.int64)
        for n, (i, j) in enumerate(zip(ii.tolist(), jj.tolist())):
            kk = key_of(f, i, j)
            v = keys.get(kk)
            if v is None:
                v = keys[kk] = len(pos)
                pos.append(point(kk))
            ids[n] = v
        # up triangles (i,j) (i+1,j) (i,j+1); down triangles (i+1,j) (i+1,j+1) (i,j+1)
        up = (ii + jj) < nu
        i0, j0 = ii[up], jj[up]
        faces.append(np.stack([ids[lat[i0, j0]], ids[lat[i0 + 1, j0]], ids[lat[i0, j0 + 1]]], 1))
        dn = (ii + jj) < nu - 1
        i1, j1 = ii[dn], jj[dn]
        faces.append(np.stack([ids[lat[i1 + 1, j1]], ids[lat[i1 + 1, j1 + 1]], ids[lat[i1, j1 + 1]]], 1))
    V = np.array(pos)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    return V, np.concatenate(faces).astype(np.int32)


def lobed_icosphere(nu=71, seed=1, n_lobes=6, amp=0.15, width=(0.3, 0.8)):
    """Perturbed icosphere (closed genus-0 surface). Returns (V, F)."""
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(n_lobes, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    widths = rng.uniform(width[0], width[1], n_lobes)
    V, F = geodesic_sphere(nu)
    theta = np.arccos(np.clip(V @ dirs.T, -1.0, 1.0))
    r = 1.0 + amp * np.exp(-theta ** 2 / (2.0 * widths[None, :] ** 2)).sum(1)
    return V * r[:, None], F


# ------------------------------------------------------------------------------------------------------------
# fused tori

TORUS_R, TORUS_r, TORUS_CX = 1.0, 0.35, (-1.8, 0.0, 1.8)
BLEND = 0.1  # fillet radius of the smooth union where two tubes meet


def tori_sdf(P):
    """Signed distance bound of the blended union of the three tori (negative inside), and its gradient. The
    union is the polynomial smooth minimum smin(a, b) = min(a, b) - k h^2 / 4, h = max(k - |a - b|, 0) / k, taken
    torus by torus: the concave creases where tubes intersect become fillets of radius ~k, so a surface
    extraction at a grid spacing well below k sees one surface sheet per cell."""
    P = np.asarray(P, np.float64)
    best = None
    for cx in TORUS_CX:
        x, y, z = P[:, 0] - cx, P[:, 1], P[:, 2]
        q = np.sqrt(x * x + y * y)
        qs = np.maximum(q, 1e-300)
        dq = q - TORUS_R
        d = np.sqrt(dq * dq + z * z)
        ds = np.maximum(d, 1e-300)
        f = d - TORUS_r
        g = np.stack([dq / ds * x / qs, dq / ds * y / qs, z / ds], 1)
        if best is None:
            best, grad = f, g
            continue
        k = BLEND
        h = np.maximum(k - np.abs(best - f), 0.0) / k
        wa = np.where(best < f, 1.0 - h / 2.0, h / 2.0)  # d smin / d best
        grad = wa[:, None] * grad + (1.0 - wa)[:, None] * g
        best = np.minimum(best, f) - k * h * h / 4.0
    return best, grad


def _project(P, steps=4):
    for _ in range(steps):
        f, g = tori_sdf(P)
        P = P - (f / np.maximum((g * g).sum(1), 1e-300))[:, None] * g
    return P


def _vertex_ring(F, nV):
    E = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    A = sp.coo_matrix((np.ones(len(E)), (E[:, 0], E[:, 1])), shape=(nV, nV)).tocsr()
    A = ((A + A.T) > 0).astype(np.float64)
    return A


def fused_tori(h=0.0219, seed=2, relax_iters=8, jitter=0.01):
    """Genus-3 surface of three fused tori. h is the grid spacing of the extraction (0.0219 gives 200,216 faces).
    Returns (V, F) with outward counter-clockwise triangles."""
    lo = np.array([TORUS_CX[0] - TORUS_R - TORUS_r - 2 * h, -TORUS_R - TORUS_r - 2 * h, -TORUS_r - 2 * h])
    hi = -lo
    n = np.ceil((hi - lo) / h).astype(int) + 1
    # grid planes shifted by 0.37 h from the bounding box, so none is tangent to the tube tops and bottoms
    # (z = +-r), where a surface touching a plane would give cells with two sheets
    xs = [lo[k] + h * (np.arange(n[k]) + 0.37) for k in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    G = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    val = tori_sdf(G)[0].reshape(n)
    inside = val < 0.0
    nx, ny, nz = n
    cell = lambda i, j, k: (i * (ny - 1) + j) * (nz - 1) + k  # noqa: E731
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    # crossing edges along each axis and the 4 cells around each
    quads = []
    sums = np.zeros((ncell, 3))
    cnt = np.zeros(ncell)
    for ax in range(3):
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[ax] = slice(0, -1)
        sl1[ax] = slice(1, None)
        a, b = inside[tuple(sl0)], inside[tuple(sl1)]
        cross = a != b
        idx = np.argwhere(cross)
        if idx.size == 0:
            continue
        i, j, k = idx[:, 0], idx[:, 1], idx[:, 2]
        v0 = val[tuple(sl0)][cross]
        v1 = val[tuple(sl1)][cross]
        t = v0 / (v0 - v1)
        p = np.stack([xs[0][i], xs[1][j], xs[2][k]], 1)
        p[:, ax] += t * h
        # the 4 cells sharing this edge: offsets in the two other axes
        o1, o2 = [d for d in range(3) if d != ax]
        cells = []
        for d1, d2 in ((0, 0), (1, 0), (1, 1), (0, 1)):
            c = np.stack([i, j, k], 1)
            c[:, o1] -= 1 - d1
            c[:, o2] -= 1 - d2
            cells.append(c)
        ok = np.ones(len(i), bool)
        for c in cells:
            ok &= (c >= 0).all(1) & (c[:, 0] < nx - 1) & (c[:, 1] < ny - 1) & (c[:, 2] < nz - 1)
        assert ok.all(), "surface reaches the grid border"
        cid = np.stack([cell(c[:, 0], c[:, 1], c[:, 2]) for c in cells], 1)
        for q in range(4):
            np.add.at(sums, cid[:, q], p)
            np.add.at(cnt, cid[:, q], 1.0)
        # orientation: the quad (c00, c10, c11, c01) winds counter-clockwise around +ax when (ax, o1, o2) is a
        # cyclic permutation of (0, 1, 2) (ax = 0 and 2, not 1); the outward normal points from the inside grid
        # point to the outside one, so the quad is reversed when the far end is inside
        flip = b[cross].copy()
        if ax == 1:
            flip = ~flip
        q4 = np.where(flip[:, None], cid[:, ::-1], cid)
        quads.append(q4)
    quads = np.concatenate(quads)
    used = np.flatnonzero(cnt > 0)
    remap = -np.ones(ncell, np.int64)
    remap[used] = np.arange(used.size)
    P = sums[used] / cnt[used, None]
    P = _project(P)
    Q = remap[quads]
    d02 = np.linalg.norm(P[Q[:, 0]] - P[Q[:, 2]], axis=1)
    d13 = np.linalg.norm(P[Q[:, 1]] - P[Q[:, 3]], axis=1)
    s = d02 <= d13
    F = np.concatenate([np.where(s[:, None], Q[:, [0, 1, 2]], Q[:, [0, 1, 3]]),
                        np.where(s[:, None], Q[:, [0, 2, 3]], Q[:, [1, 2, 3]])]).astype(np.int64)
    F = _check_manifold_and_fix(F, P.shape[0])
    # tangential relaxation with projection
    A = _vertex_ring(F, P.shape[0])
    deg = np.asarray(A.sum(1)).ravel()
    for _ in range(relax_iters):
        c = (A @ P) / deg[:, None]
        nrm = tori_sdf(P)[1]
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        d = c - P
        d -= (d * nrm).sum(1, keepdims=True) * nrm
        P = _project(P + 0.8 * d)
    rng = np.random.default_rng(seed)
    el = mean_edge_length(P, F)
    P = P + rng.normal(size=P.shape) * (jitter * el)
    return P, F.astype(np.int32)


def _check_manifold_and_fix(F, nV):
    """Every edge in exactly two triangles with opposite directions; raises otherwise."""
    E = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    key = E[:, 0] * nV + E[:, 1]
    if np.unique(key).size != key.size:
        raise ValueError("surface nets: a directed edge appears twice (non-manifold cell)")
    rkey = E[:, 1] * nV + E[:, 0]
    if not np.isin(rkey, key).all():
        raise ValueError("surface nets: open edge")
    return F


# ------------------------------------------------------------------------------------------------------------
# topology helpers and the cut

def mean_edge_length(V, F):
    E = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    return float(np.linalg.norm(V[E[:, 0]] - V[E[:, 1]], axis=1).mean())


def edges_of(F):
    """Unique undirected edges (E, 2) sorted, and per triangle corner-edge the edge id (F, 3): edge k of triangle
    t is (F[t,k], F[t,(k+1)%3])."""
    F = np.asarray(F, np.int64)
    nV = int(F.max()) + 1
    he = np.stack([F, np.roll(F, -1, axis=1)], 2).reshape(-1, 2)
    lo, hi = np.minimum(he[:, 0], he[:, 1]), np.maximum(he[:, 0], he[:, 1])
    key = lo * nV + hi
    uk, inv = np.unique(key, return_inverse=True)
    E = np.stack([uk // nV, uk % nV], 1)
    return E, inv.reshape(-1, 3)


def euler_characteristic(F):
    F = np.asarray(F)
    E, _ = edges_of(F)
    return int(np.unique(F).size - E.shape[0] + F.shape[0])


def genus_closed(F):
    return (2 - euler_characteristic(F)) // 2


def farthest_point_path(V, F):
    """TriMesh::farthestPointCut's path (TriMesh.cpp:1250-1278): vertex ids from p (farthest from F[0,0]) to
    the vertex farthest from p, along shortest paths of the rest edge-length graph."""
    E, _ = edges_of(F)
    nV = V.shape[0]
    w = np.linalg.norm(V[E[:, 0]] - V[E[:, 1]], axis=1)
    G = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([E[:, 0], E[:, 1]]),
                                                np.concatenate([E[:, 1], E[:, 0]]))), shape=(nV, nV)).tocsr()
    d0 = csg.dijkstra(G, indices=int(F[0, 0]))
    p = int(np.argmax(d0))
    d1, pred = csg.dijkstra(G, indices=p, return_predecessors=True)
    q = int(np.argmax(d1))
    path = [q]
    while path[-1] != p:
        path.append(int(pred[path[-1]]))
    return np.array(path[::-1], np.int64)


def greedy_loops_cut(V, F, root=None):
    """Edges of a cut graph that opens a closed surface into a disc: the greedy shortest system of loops
    (Erickson & Whittlesey, SODA 2005). T = the shortest-path tree of the rest edge lengths from the root vertex
    (F[0, 0]); every edge e = (u, v) not in T closes the loop d(u) + |e| + d(v); the dual graph's maximum
    spanning tree over the edges not in T, weighted by those loop lengths, leaves 2g edges whose loops (through
    the root) form the cut graph: the edges no dual tree edge crosses, dangling branches pruned. Returns
    (edge ids, E)."""
    F = np.asarray(F, np.int64)
    E, te = edges_of(F)
    nF = F.shape[0]
    nV = V.shape[0]
    w = np.linalg.norm(V[E[:, 0]] - V[E[:, 1]], axis=1)
    G = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([E[:, 0], E[:, 1]]),
                                                np.concatenate([E[:, 1], E[:, 0]]))), shape=(nV, nV)).tocsr()
    r = int(F[0, 0]) if root is None else int(root)
    dist, pred = csg.dijkstra(G, indices=r, return_predecessors=True)
    in_T = np.zeros(E.shape[0], bool)
    child = np.flatnonzero(pred >= 0)
    key = E[:, 0] * nV + E[:, 1]
    a, b = np.minimum(child, pred[child]), np.maximum(child, pred[child])
    in_T[np.searchsorted(key, a * nV + b)] = True
    loop_len = dist[E[:, 0]] + dist[E[:, 1]] + w
    # dual edges: the two triangles of every edge
    owners = np.full((E.shape[0], 2), -1, np.int64)
    flat = te.ravel()
    order = np.argsort(flat, kind="stable")
    tri = order // 3
    first = np.ones(order.size, bool)
    first[1:] = flat[order][1:] != flat[order][:-1]
    owners[flat[order][first], 0] = tri[first]
    owners[flat[order][~first], 1] = tri[~first]
    assert (owners >= 0).all(), "not a closed manifold surface"
    ne = np.flatnonzero(~in_T)
    # maximum spanning tree = minimum spanning tree of (max + 1 - weight) (weights stay positive)
    ww = loop_len[ne].max() + 1.0 - loop_len[ne]
    D = sp.coo_matrix((ww, (owners[ne, 0], owners[ne, 1])), shape=(nF, nF)).tocsr()
    M = csg.minimum_spanning_tree(D).tocoo()
    assert M.nnz == nF - 1, "dual graph not connected"
    # map the tree's dual edges back to mesh edges (a pair of triangles shares one edge)
    dkey = {}
    for k, e in enumerate(ne.tolist()):
        t1, t2 = int(owners[e, 0]), int(owners[e, 1])
        dkey[(min(t1, t2), max(t1, t2))] = e
    crossed = np.zeros(E.shape[0], bool)
    for t1, t2 in zip(M.row.tolist(), M.col.tolist()):
        crossed[dkey[(min(t1, t2), max(t1, t2))]] = True
    cut = ~crossed
    # prune dangling branches (vertices of degree 1 in the cut graph)
    while True:
        ce = np.flatnonzero(cut)
        deg = np.bincount(E[ce].ravel(), minlength=nV)
        leaf = deg == 1
        drop = ce[leaf[E[ce, 0]] | leaf[E[ce, 1]]]
        if drop.size == 0:
            break
        cut[drop] = False
    return np.flatnonzero(cut), E


def tree_cotree_cut(V, F):
    """Edges of a cut graph that opens a closed surface into a disc: the edges no dual spanning-tree edge crosses
    (breadth-first over triangles from triangle 0), with dangling branches pruned. On the fused tori its total
    length is about two thirds of the greedy system of loops' (every greedy loop returns to one root), and the
    Tutte map it leads to is about half as distorted. Returns (edge ids, E)."""
    F = np.asarray(F, np.int64)
    E, te = edges_of(F)
    nF = F.shape[0]
    # dual graph: triangles sharing an edge
    owners = np.full((E.shape[0], 2), -1, np.int64)
    flat = te.ravel()
    order = np.argsort(flat, kind="stable")
    tri = (order // 3)
    first = np.ones(order.size, bool)
    first[1:] = flat[order][1:] != flat[order][:-1]
    owners[flat[order][first], 0] = tri[first]
    owners[flat[order][~first], 1] = tri[~first]
    assert (owners >= 0).all(), "not a closed manifold surface"
    D = sp.coo_matrix((np.arange(E.shape[0]) + 1, (owners[:, 0], owners[:, 1])), shape=(nF, nF)).tocsr()
    D = D + D.T
    _, pred = csg.breadth_first_order(D, 0, directed=False, return_predecessors=True)
    tree_edge = np.zeros(E.shape[0], bool)
    t = np.flatnonzero(pred >= 0)
    tree_edge[np.asarray(D[t, pred[t]]).ravel().astype(np.int64) - 1] = True
    cut = ~tree_edge
    # prune dangling branches (vertices of degree 1 in the cut graph)
    nV = V.shape[0]
    while True:
        ce = np.flatnonzero(cut)
        deg = np.bincount(E[ce].ravel(), minlength=nV)
        leaf = deg == 1
        drop = ce[leaf[E[ce, 0]] | leaf[E[ce, 1]]]
        if drop.size == 0:
            break
        cut[drop] = False
    return np.flatnonzero(cut), E


def path_edges(path, F):
    E, _ = edges_of(F)
    nV = int(np.asarray(F).max()) + 1
    key = E[:, 0] * nV + E[:, 1]
    a, b = np.minimum(path[:-1], path[1:]), np.maximum(path[:-1], path[1:])
    pos = np.searchsorted(key, a * nV + b)
    assert (key[pos] == a * nV + b).all(), "path does not follow mesh edges"
    return pos, E


def cut_along(V, F, cut_edge_ids):
    """Split the mesh along the given edges. Returns (Vcut, Fcut, cohE, vmap) where vmap[v'] is the uncut vertex
    of copy v' and cohE rows are [a', b', a'', b''] (copies of edge (a, b) on its two sides)."""
    F = np.asarray(F, np.int64)
    nF = F.shape[0]
    E, te = edges_of(F)
    is_cut = np.zeros(E.shape[0], bool)
    is_cut[cut_edge_ids] = True
    # corners c = 3 t + k; across an uncut edge (a, b) of triangles t1 (a -> b) and t2 (b -> a) join the corners of
    # a in both and of b in both
    he_t = np.repeat(np.arange(nF), 3)
    he_k = np.tile(np.arange(3), nF)
    eid = te.ravel()
    order = np.lexsort((he_t, eid))
    e_sorted = eid[order]
    pair_start = np.flatnonzero(np.r_[True, e_sorted[1:] != e_sorted[:-1]])
    assert (np.diff(np.r_[pair_start, e_sorted.size]) == 2).all(), "every edge must have two triangles"
    h1, h2 = order[pair_start], order[pair_start + 1]
    e = e_sorted[pair_start]
    t1, k1, t2, k2 = he_t[h1], he_k[h1], he_t[h2], he_k[h2]
    a1, b1 = 3 * t1 + k1, 3 * t1 + (k1 + 1) % 3          # corners of the edge's start and end in t1
    a2, b2 = 3 * t2 + (k2 + 1) % 3, 3 * t2 + k2          # the same vertices in t2 (opposite direction)
    keep = ~is_cut[e]
    rows = np.concatenate([a1[keep], b1[keep]])
    cols = np.concatenate([a2[keep], b2[keep]])
    G = sp.coo_matrix((np.ones(rows.size), (rows, cols)), shape=(3 * nF, 3 * nF))
    ncomp, lab = csg.connected_components(G, directed=False)
    # number copies by the lowest corner of each class, so ids follow the triangle order
    first_corner = np.full(ncomp, 3 * nF, np.int64)
    np.minimum.at(first_corner, lab, np.arange(3 * nF))
    rank = np.empty(ncomp, np.int64)
    rank[np.argsort(first_corner)] = np.arange(ncomp)
    cid = rank[lab]
    Fcut = cid.reshape(nF, 3)
    vmap = np.empty(ncomp, np.int64)
    vmap[cid] = F.ravel()
    Vcut = V[vmap]
    cc = ~keep
    cohE = np.stack([cid[a1[cc]], cid[b1[cc]], cid[a2[cc]], cid[b2[cc]]], 1)
    return Vcut, Fcut.astype(np.int32), cohE.astype(np.int32), vmap


def boundary_loops(F):
    """Boundary loops of a mesh with counter-clockwise triangles (mesh on the left), lowest-start first."""
    F = np.asarray(F, np.int64)
    nV = int(F.max()) + 1
    he = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    key = he[:, 0] * nV + he[:, 1]
    rev = he[:, 1] * nV + he[:, 0]
    bnd = he[~np.isin(key, rev)]
    nxt = dict(zip(bnd[:, 0].tolist(), bnd[:, 1].tolist()))
    loops, seen = [], set()
    for s in sorted(nxt):
        if s in seen:
            continue
        loop = [s]
        seen.add(s)
        while nxt[loop[-1]] != s:
            loop.append(nxt[loop[-1]])
            seen.add(loop[-1])
        loops.append(np.array(loop, np.int32))
    return loops


def cut_to_disc(V, F):
    """Cut a closed surface to a disc (genus 0: farthest-point path; otherwise tree-cotree). Returns
    (Vcut, Fcut, cohE, vmap)."""
    g = genus_closed(F)
    if g == 0:
        path = farthest_point_path(V, F)
        ids, _ = path_edges(path, F)
    else:
        ids, _ = tree_cotree_cut(V, F)
    out = cut_along(V, F, ids)
    assert euler_characteristic(out[1]) == 1 and len(boundary_loops(out[1])) == 1, "the cut is not a disc"
    return out


def workload(name):
    """The closed-surface workloads, cut to a disc: (Vcut, Fcut, cohE, meta)."""
    if name == "icosphere_100k":
        V, F = lobed_icosphere(71, seed=1, width=(0.3, 0.8))
    elif name == "icosphere_1m":
        V, F = lobed_icosphere(224, seed=4, width=(0.2, 0.6))
    elif name == "tori_200k":
        V, F = fused_tori(seed=2)
    elif name.startswith("icosphere_nu"):
        V, F = lobed_icosphere(int(name[len("icosphere_nu"):]), seed=1, width=(0.3, 0.8))
    elif name.startswith("tori_h"):
        V, F = fused_tori(h=float(name[len("tori_h"):]), seed=2)
    else:
        raise KeyError(name)
    Vc, Fc, cohE, vmap = cut_to_disc(V, F)
    return Vc, Fc, cohE, {"faces": int(F.shape[0]), "vertices": int(V.shape[0]), "genus": genus_closed(F),
                          "cut_vertices": int(Vc.shape[0]), "seam_edges": int(cohE.shape[0])}
