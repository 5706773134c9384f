"""Pins of the oracle's scaffold composition (tests/_helpers.py::ScaffoldOracle) that do not reuse its parts:
central differences of the combined objective, the isometry closed form, the fixed-vertex diagonal by hand, and
the reference's triplet order written out by hand for a two-triangle mesh.

The composition restates Optimizer::computeEnergyVal/Gradient/Hessian (Optimizer.cpp:764-843) with
Scaffold::augmentGradient / augmentProxyMatrix (Scaffold.cpp:210-248): a0 E_SD(mesh) + (w_scaf / F_air) E_SD(air,
uniform weights), air vertex l at global id localVI2Global[l] (boundary air vertices are mesh vertices, the others
follow the mesh, Scaffold.cpp:179-184), fixed mesh vertices fixed in the air mesh too (:186-192).
"""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd.meshgen import height_field_disc

from _helpers import ScaffoldOracle, distort, grid_air_mesh


def _setup(m=6, seed=2, a0=0.7, w_scaf=0.05, flat=False, distort_air=True):
    V, F, UV = height_field_disc(m, seed=seed)
    if flat:
        V = V.copy()
        V[:, 2] = 0.0
    airUV, Fair, bnd = grid_air_mesh(m, rings=2)
    so = ScaffoldOracle(V, F, airUV, Fair, bnd, 0.0, a0, w_scaf)
    X = np.concatenate([UV, airUV[len(bnd):]])
    if distort_air:
        X = distort(X, seed + 10, amp=0.02)
    return V, F, airUV, Fair, bnd, so, X


def _free_dofs(so, X):
    nVt = X.shape[0]
    fixed = np.zeros(nVt, bool)
    fixed[0] = True  # vertex 0 (TriMesh.cpp:325-328), also fixed in the air mesh
    return np.flatnonzero(~np.repeat(fixed, 2))


def test_scaffold_gradient_matches_central_differences():
    V, F, airUV, Fair, bnd, so, X = _setup()
    assert O.check_inversion(F, X[: V.shape[0]])
    g = so.gradient(X)
    # fixed rows are zero (SymDirichletEnergy.cpp:300-303 in both blocks)
    assert g[0] == 0.0 and g[1] == 0.0
    h = 1e-6
    x = X.reshape(-1)
    for k in _free_dofs(so, X):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd = (so.energy(xp.reshape(-1, 2))[0] - so.energy(xm.reshape(-1, 2))[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, np.abs(g).max()), (k, fd, g[k])


def test_scaffold_hessian_matches_central_differences_of_gradient():
    V, F, airUV, Fair, bnd, so, X = _setup()
    ia, ja, a = so.hessian_upper(X, project=False)
    n = 2 * X.shape[0]
    H = np.zeros((n, n))
    for r in range(n):
        for q in range(ia[r], ia[r + 1]):
            H[r, ja[q]] = a[q]
            H[ja[q], r] = a[q]
    free = _free_dofs(so, X)
    h = 1e-6
    x = X.reshape(-1)
    FD = np.zeros((n, n))
    for k in free:
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        FD[:, k] = (so.gradient(xp.reshape(-1, 2)) - so.gradient(xm.reshape(-1, 2))) / (2 * h)
    Hf, FDf = H[np.ix_(free, free)], FD[np.ix_(free, free)]
    assert np.abs(Hf - FDf).max() <= 1e-5 * np.abs(Hf).max()
    # the air block really contributes: the mesh-only Hessian differs on boundary rows
    Hm = O.element_hessians(F, X[: V.shape[0]], so.feat, project=False)
    assert np.abs(Hf).max() > 0 and Hm.shape[0] == F.shape[0]


def test_scaffold_energy_closed_form_at_isometry():
    """Flat rest mesh with UV = its rest xy, air mesh whose rest shape is its own UV (nothing clamped): every mesh
    triangle contributes 4 w_t (sum w = 1) and every air triangle 4, so E = 4 a0 + (w_scaf / F_air) 4 F_air."""
    a0, w_scaf = 0.37, 0.021
    V, F, airUV, Fair, bnd, so, X = _setup(m=7, a0=a0, w_scaf=w_scaf, flat=True, distort_air=False)
    E, Em, Ea = so.energy(X)
    assert Em == pytest.approx(4.0, rel=1e-13)
    assert Ea == pytest.approx(4.0 * w_scaf, rel=1e-13)
    assert E == pytest.approx(4.0 * a0 + 4.0 * w_scaf, rel=1e-13)
    # and it is a stationary point: the gradient vanishes to rounding
    assert np.abs(so.gradient(X)).max() <= 1e-12


def test_scaffold_fixed_vertex_diagonal_by_hand():
    """Vertex 0 is fixed in the mesh and, being on the boundary, in the air mesh: its rows hold only the
    diagonal, a0 (the mesh's identity triplet scaled by energyParams[0]) + w_scaf / F_air (the air mesh's
    identity triplet scaled by the scaffold weight)."""
    a0, w_scaf = 0.61, 0.043
    V, F, airUV, Fair, bnd, so, X = _setup(a0=a0, w_scaf=w_scaf)
    assert 0 in list(bnd)
    ia, ja, a = so.hessian_upper(X)
    nFair = Fair.shape[0]
    for r in (0, 1):
        assert ia[r + 1] - ia[r] == 1 and ja[ia[r]] == r
        assert a[ia[r]] == a0 + w_scaf / nFair
    # a free boundary vertex's diagonal mixes both blocks: it is larger than the mesh's own
    v = int(bnd[1])
    Hm = O.element_hessians(F, X[: V.shape[0]], so.feat)
    dm = sum(a0 * Hm[t][2 * list(F[t]).index(v), 2 * list(F[t]).index(v)] for t in range(F.shape[0]) if v in F[t])
    assert a[ia[2 * v]] > dm


# Reference triplet order of SymDirichletEnergy::computeHessian(V, I, J) for F = [(0,1,2), (2,1,3)] with vertex 0
# fixed, written out by hand from IglUtils::addBlockToMatrix (IglUtils.cpp:396-451: triangles in order, local
# vertex pairs (i, j) over free vertices, then dimI, dimJ) and addDiagonalToMatrix (:361-394, fixed vertices
# ascending, last).
TWO_TRI_IJ = [
    # triangle 0, vertices (0*, 1, 2): blocks (1,1) (1,2) (2,1) (2,2)
    (2, 2), (2, 3), (3, 2), (3, 3), (2, 4), (2, 5), (3, 4), (3, 5),
    (4, 2), (4, 3), (5, 2), (5, 3), (4, 4), (4, 5), (5, 4), (5, 5),
    # triangle 1, vertices (2, 1, 3): blocks (2,2) (2,1) (2,3) (1,2) (1,1) (1,3) (3,2) (3,1) (3,3)
    (4, 4), (4, 5), (5, 4), (5, 5), (4, 2), (4, 3), (5, 2), (5, 3), (4, 6), (4, 7), (5, 6), (5, 7),
    (2, 4), (2, 5), (3, 4), (3, 5), (2, 2), (2, 3), (3, 2), (3, 3), (2, 6), (2, 7), (3, 6), (3, 7),
    (6, 4), (6, 5), (7, 4), (7, 5), (6, 2), (6, 3), (7, 2), (7, 3), (6, 6), (6, 7), (7, 6), (7, 7),
    # identity on the fixed vertex's DOFs
    (0, 0), (1, 1),
]


def two_triangle_case():
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.2]])
    F = np.array([[0, 1, 2], [2, 1, 3]], np.int32)
    UV = np.array([[0.0, 0.0], [1.1, 0.05], [-0.1, 0.9], [1.05, 1.2]])
    return V, F, UV


def check_two_triangle_triplets(I, J, Vv, H):
    """H: the two projected element Hessians (2, 6, 6)."""
    assert list(zip(I.tolist(), J.tolist())) == TWO_TRI_IJ
    # values by position: entry 5 is triangle 0's (vertex 1, x) x (vertex 2, y) = H0[2, 5]; entry 26 is triangle
    # 1's (local 0 = vertex 2, y) x (local 2 = vertex 3, x) = H1[1, 4]; entry 20 is triangle 1's (vertex 2, x) x
    # (vertex 1, x) = H1[0, 2]
    assert Vv[5] == H[0][2, 5]
    assert Vv[26] == H[1][1, 4]
    assert Vv[20] == H[1][0, 2]
    assert Vv[52] == 1.0 and Vv[53] == 1.0


def test_oracle_triplet_order_two_triangles_by_hand():
    V, F, UV = two_triangle_case()
    feat = O.RestFeatures(V, F)
    H = O.element_hessians(F, UV, feat)
    I, J, Vv = O.hessian_triplets(F, H, fixed=(0,))
    check_two_triangle_triplets(I, J, Vv, H)
