"""GPU Tutte initial map (oc_tutte) against the reference's libigl path.

Tolerances: UV to 1e-10 absolute against the reference's direct solve at PCG
tolerance 1e-13 (the linear system is well scaled: entries are small integers,
boundary values on the unit circle); the boundary loop is integer data and is
checked through its length and the mapped positions; at 200k faces the
harmonic residual, the circle and triangle orientations are checked directly.
"""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd import capi
from optcuts_amd.meshgen import height_field_disc

pytestmark = pytest.mark.gpu

CASES = [(n, l) for n in ("tutte11", "tutte21", "tutte21h", "tutte3c") for l in (0, 1)]


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("precond", [capi.PRECOND_JACOBI, capi.PRECOND_AMG])
@pytest.mark.parametrize("name,layout", CASES)
def test_tutte_matches_reference_golden(ctx, golden, name, layout, precond):
    V, F = golden[name + "_Vrest"], golden[name + "_F"]
    key = "%s_l%d" % (name, layout)
    ctx.set_mesh(V, F, fixed=(0,))
    W, nl, nb, it = ctx.tutte(layout=layout, precond=precond, rel_tol=1e-13, max_iter=20000)
    print(key, precond, nl, nb, it, flush=True)
    assert nl == len(golden[key + "_loops"])
    assert nb == len(golden[key + "_bnd"])
    np.testing.assert_allclose(W, golden[key + "_UV"], rtol=0, atol=1e-10)
    b = golden[key + "_bnd"]
    np.testing.assert_allclose(W[b], golden[key + "_UV"][b], rtol=0, atol=1e-15)
    np.testing.assert_allclose(ctx.get_uv(), W, rtol=0, atol=0)


def test_tutte_then_newton_matches_oracle(ctx, golden):
    """Newton from the Tutte map follows the oracle's trajectory from the reference's map."""
    V, F = golden["tutte21_Vrest"], golden["tutte21_F"]
    UV0 = golden["tutte21_l0_UV"]
    nV = V.shape[0]
    rel = 1e-16 * nV / (nV - 1)
    tr, _, _ = O.newton_run(V, F, UV0, a0=1.0, max_iter=10, rel_gl2_tol=rel, allow_edec_rel_tol=False)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.tutte(precond=capi.PRECOND_AMG, rel_tol=1e-14, max_iter=20000)
    E0 = ctx.newton_init(a0=1.0, rel_gl2_tol=rel, allow_edec_rel_tol=False, pcg_rel_tol=1e-12,
                         pcg_max_iter=20000)
    assert E0 == pytest.approx(tr[0, 0], rel=1e-9)
    for k in range(1, len(tr)):
        s, done = ctx.newton_step()
        assert s["E_total"] == pytest.approx(tr[k, 0], rel=1e-9), k
        if done:
            break


def test_tutte_rejects_bad_topology(ctx):
    V, F, _ = height_field_disc(5, seed=0)
    ctx.set_mesh(np.vstack([V, [[2.0, 2.0, 0.0]]]), F, fixed=(0,))
    with pytest.raises(RuntimeError, match="no triangle"):
        ctx.tutte()
    # closed surface: a tetrahedron
    Vt = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    Ft = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]], np.int32)
    ctx.set_mesh(Vt, Ft, fixed=(0,))
    with pytest.raises(RuntimeError, match="component 0 has no boundary"):
        ctx.tutte()
    # an edge held by three triangles
    Vn = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1.0]])
    Fn = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.int32)
    ctx.set_mesh(Vn, Fn, fixed=(0,))
    with pytest.raises(RuntimeError, match="more than two triangles"):
        ctx.tutte()


def test_tutte_200k_faces_properties(ctx):
    """Full-size workload (317x317 grid, 199,712 faces): AMG and Jacobi agree, the
    boundary lies on the circle, interior vertices average their neighbours and every
    triangle keeps its orientation."""
    V, F, _ = height_field_disc(317, seed=0)
    ctx.set_mesh(V, F, fixed=(0,))
    W, nl, nb, it_amg = ctx.tutte(precond=capi.PRECOND_AMG, rel_tol=1e-12, max_iter=5000)
    Wj, _, nbj, it_jac = ctx.tutte(precond=capi.PRECOND_JACOBI, rel_tol=1e-12, max_iter=200000)
    print("amg", it_amg, "jacobi", it_jac, flush=True)
    assert nl == 1 and nb == nbj == 4 * 316
    assert it_amg < 100
    assert np.abs(W - Wj).max() < 1e-8
    r = np.linalg.norm(W, axis=1)
    onc = np.abs(r - 1.0) < 1e-14
    assert onc.sum() >= nb
    # uniform-Laplacian residual at interior vertices (grid interior: every edge in two triangles)
    nV = V.shape[0]
    deg = np.zeros(nV)
    acc = np.zeros((nV, 2))
    for a, b in ((0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2)):
        np.add.at(acc, F[:, a], W[F[:, b]])
        np.add.at(deg, F[:, a], 1.0)
    interior = np.ones(nV, bool)
    m = 317
    j, i = np.divmod(np.arange(nV), m)
    interior &= (i > 0) & (i < m - 1) & (j > 0) & (j < m - 1)
    res = np.abs(acc[interior] - deg[interior, None] * W[interior]).max()
    assert res < 1e-9
    assert O.check_inversion(F, W)
