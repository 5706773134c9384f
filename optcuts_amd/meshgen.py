"""Deterministic synthetic meshes for the geometry-step workloads.

Disc height field (the reference's DistMin input class): an m x m vertex grid
on [0,1]^2, each square split along a diagonal whose direction is drawn
Bernoulli(0.5), heights z = sum of 8 Gaussian bumps with centres U[0,1]^2,
amplitudes U(0.1,0.4) and widths U(0.05,0.2); the initial UV is the xy
projection, so every face is counter-clockwise in UV. Draws come from
numpy.random.default_rng(seed) in that order (flips, centres, amplitudes,
widths).

m = 51 gives the 5,000-face case; m = 317 gives 199,712 faces, the 200k-face
workload. Vertex 0 is the (0,0) corner, the vertex the reference fixes
(TriMesh.cpp:325-328).
"""
import numpy as np


def height_field_disc(m=51, seed=0, n_bumps=8):
    """Return (Vrest (V,3) f64, F (F,3) int32, UV (V,2) f64)."""
    rng = np.random.default_rng(seed)
    s = m - 1
    flips = rng.random(s * s) < 0.5
    centres = rng.random((n_bumps, 2))
    amps = rng.uniform(0.1, 0.4, n_bumps)
    sigmas = rng.uniform(0.05, 0.2, n_bumps)

    xs = np.linspace(0.0, 1.0, m)
    X, Y = np.meshgrid(xs, xs, indexing="xy")  # row j = y index, col i = x index
    P = np.stack([X.ravel(), Y.ravel()], axis=1)
    d2 = ((P[:, None, :] - centres[None, :, :]) ** 2).sum(-1)
    Z = (amps[None, :] * np.exp(-d2 / (2.0 * sigmas[None, :] ** 2))).sum(-1)
    Vrest = np.stack([P[:, 0], P[:, 1], Z], axis=1)

    j, i = np.meshgrid(np.arange(s), np.arange(s), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    v00 = j * m + i
    v10 = v00 + 1
    v01 = v00 + m
    v11 = v01 + 1
    fl = flips
    # diagonal v00-v11 (not flipped) or v10-v01 (flipped); all CCW in xy
    t1 = np.where(fl[:, None], np.stack([v00, v10, v01], 1), np.stack([v00, v10, v11], 1))
    t2 = np.where(fl[:, None], np.stack([v10, v11, v01], 1), np.stack([v00, v11, v01], 1))
    F = np.empty((2 * s * s, 3), dtype=np.int32)
    F[0::2] = t1
    F[1::2] = t2
    UV = P.copy()
    return Vrest, F, UV


def faces_for_target(target_faces):
    """Grid size m whose height-field disc has about target_faces faces."""
    return int(round(np.sqrt(target_faces / 2.0))) + 1


def batched_disc_charts(n_charts=64, lo=500, hi=20000, base_seed=1000):
    """Charts of uneven size: F_i = exp(U(ln lo, ln hi)), each a height-field disc.

    Sizes come from default_rng(base_seed); chart i uses seed base_seed + i.
    Returns a list of (Vrest, F, UV).
    """
    rng = np.random.default_rng(base_seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_charts))
    charts = []
    for k, fk in enumerate(sizes):
        m = max(3, int(round(np.sqrt(fk / 2.0))) + 1)
        charts.append(height_field_disc(m, seed=base_seed + k))
    return charts


def union_of_charts(charts):
    """Concatenate charts into one mesh for batched solving.

    Returns (V, F, UV, vertexStart, faceStart, fixed): chart k owns vertices
    [vertexStart[k], vertexStart[k+1]) and triangles [faceStart[k],
    faceStart[k+1]); fixed lists each chart's vertex 0 (the reference fixes
    vertex 0 of a mesh, TriMesh.cpp:325-328).
    """
    vs, fs = [0], [0]
    Vs, Fs, UVs, fixed = [], [], [], []
    for V, F, UV in charts:
        off = vs[-1]
        Vs.append(V)
        Fs.append(F + off)
        UVs.append(UV)
        fixed.append(off)
        vs.append(off + V.shape[0])
        fs.append(fs[-1] + F.shape[0])
    return (np.concatenate(Vs), np.concatenate(Fs).astype(np.int32), np.concatenate(UVs),
            np.array(vs, np.int32), np.array(fs, np.int32), np.array(fixed, np.int32))


def boundary_loop_ccw(F):
    """The boundary loop of a disc mesh with counter-clockwise triangles, in
    the order that keeps the mesh on its left (vertex ids, int32)."""
    F = np.asarray(F)
    he = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    key = he[:, 0].astype(np.int64) * (int(F.max()) + 1) + he[:, 1]
    rev = he[:, 1].astype(np.int64) * (int(F.max()) + 1) + he[:, 0]
    bnd = he[~np.isin(key, rev)]
    nxt = dict(zip(bnd[:, 0].tolist(), bnd[:, 1].tolist()))
    start = int(bnd[:, 0].min())
    loop = [start]
    while nxt[loop[-1]] != start:
        loop.append(nxt[loop[-1]])
    return np.array(loop, np.int32)


def band_air_mesh(UV, loop, rings=2, seed=0, spacing=None):
    """Air mesh for the bijectivity scaffold: `rings` rows of quads outside the
    boundary loop (mesh on its left) at the current UV, rows offset along the
    outward vertex normals by multiples of the mean boundary edge length (or
    `spacing`), each quad split along a diagonal drawn from default_rng(seed)
    (a fresh triangulation per call, like a re-meshed air region). Returns
    (airUV, Fair, bnd) in the layout of the scaffold interface: the first
    len(loop) air vertices are the boundary vertices `loop`. The spacing halves
    until every air triangle is counter-clockwise; ValueError if none works."""
    P = np.asarray(UV, np.float64)[loop]
    nb = len(loop)
    t = np.roll(P, -1, 0) - np.roll(P, 1, 0)
    nrm = np.stack([t[:, 1], -t[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    h = spacing if spacing else float(np.linalg.norm(np.roll(P, -1, 0) - P, axis=1).mean())
    flip = np.random.default_rng(seed).random((rings, nb)) < 0.5
    i = np.arange(nb)
    j = (i + 1) % nb
    tris = []
    for r in range(rings):
        a, b, c, d = r * nb + i, r * nb + j, (r + 1) * nb + i, (r + 1) * nb + j
        t1 = np.where(flip[r][:, None], np.stack([a, c, b], 1), np.stack([a, c, d], 1))
        t2 = np.where(flip[r][:, None], np.stack([b, c, d], 1), np.stack([a, d, b], 1))
        tris += [t1, t2]
    Fair = np.concatenate(tris).astype(np.int32)
    for _ in range(10):
        airUV = np.concatenate([P] + [P + r * h * nrm for r in range(1, rings + 1)])
        e1 = airUV[Fair[:, 1]] - airUV[Fair[:, 0]]
        e2 = airUV[Fair[:, 2]] - airUV[Fair[:, 0]]
        if ((e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]) > 0).all():
            return airUV, Fair, np.asarray(loop, np.int32)
        h *= 0.5
    raise ValueError("band_air_mesh: no spacing gives a counter-clockwise band")
