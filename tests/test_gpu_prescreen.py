"""GPU split-query pre-screen (TriMesh::querySplit's candidate filter and
SymDirichletEnergy::computeDivGradPerVert) against the oracle.

Tolerances: keys 1e-12 relative (the element pass contracts multiply-adds, the
reference rounds each operation); candidate lists equal (the distorted maps
used here have no keys that close together)."""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd import capi
from optcuts_amd.meshgen import height_field_disc

from _helpers import distort

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("m,seed", [(51, 0), (31, 3)])
def test_div_grad_matches_oracle(ctx, m, seed):
    V, F, UV = height_field_disc(m, seed=seed)
    UV = distort(UV, seed, amp=0.05)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    got = ctx.div_grad_per_vert()
    want = O.div_grad_per_vert(F, UV, O.RestFeatures(V, F))
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    np.testing.assert_array_equal(got == 0, want == 0)


@pytest.mark.parametrize("interior", [False, True])
def test_split_candidates_match_oracle(ctx, interior):
    m = 51
    V, F, UV = height_field_disc(m, seed=1)
    UV = distort(UV, 1, amp=0.05)
    nV = V.shape[0]
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    feat = O.RestFeatures(V, F)
    dg = O.div_grad_per_vert(F, UV, feat)
    w = np.random.default_rng(5).uniform(0.5, 2.0, nV)
    for weight in (None, w):
        got, ns = ctx.split_candidates(interior, 0.6, weight)
        want, wns = O.split_candidates(F, nV, dg, weight, interior, 0.6)
        print(interior, ns, len(got), flush=True)
        assert ns == wns and len(got) == len(want) >= 2
        np.testing.assert_array_equal(got, want)


def test_split_candidates_after_newton_steps(ctx):
    """Keys at the UV the device-resident Newton driver reached (the state the topology step queries)."""
    V, F, UV = height_field_disc(41, seed=2)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG,
                    pcg_rel_tol=1e-8, pcg_max_iter=1000)
    for _ in range(4):
        ctx.newton_step()
    UVn = ctx.get_uv()
    got = ctx.div_grad_per_vert()
    want = O.div_grad_per_vert(F, UVn, O.RestFeatures(V, F))
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    c_got, ns = ctx.split_candidates(False)
    c_want, wns = O.split_candidates(F, V.shape[0], want)
    assert ns == wns and list(c_got) == list(c_want)


def test_split_candidates_200k_faces(ctx):
    """Full size: 1,264 boundary vertices; those with more than two neighbours are the map, its first
    int(size^0.8) the candidates; interior: the (317-4)^2 vertices away from the boundary, int(size^0.6)."""
    m = 317
    V, F, UV = height_field_disc(m, seed=0)
    UV = distort(UV, 9, amp=0.03)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    deg = np.bincount(F.ravel(), minlength=V.shape[0])
    corners_single = int((deg[[0, m - 1, m * (m - 1), m * m - 1]] == 1).sum())
    cand, ns = ctx.split_candidates(False)
    assert ns == 4 * (m - 1) - corners_single
    assert len(cand) == int(ns ** 0.8) and len(set(cand)) == len(cand)
    dg = ctx.div_grad_per_vert()
    assert np.all(np.diff(-dg[cand]) >= 0)  # key order
    cand_in, ns_in = ctx.split_candidates(True, 0.6)
    assert ns_in == (m - 4) ** 2 and len(cand_in) == int(ns_in ** 0.6)
    # the host selection (filters, map order, cut; chunk-sorted and merged at this size) against the oracle's
    # filter applied to the same GPU keys
    for interior, got in ((False, cand), (True, cand_in)):
        want, wns = O.split_candidates(F, V.shape[0], dg, None, interior, 0.6)
        assert wns == (ns_in if interior else ns) and list(got) == list(want)


def test_seam_sparsity_matches_oracle(ctx):
    """TriMesh::computeSeamSparsity on the GPU: random cohesive edges over a 5k-face disc (some boundary rows
    with -1, some whose sides coincide in UV), both modes, against the oracle (1e-12; the GPU sums in a
    different order)."""
    V, F, UV = height_field_disc(51, seed=2)
    UV = distort(UV, 2, amp=0.04)
    nV = V.shape[0]
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    feat = O.RestFeatures(V, F)
    rng = np.random.default_rng(3)
    coh = rng.integers(0, nV, size=(3000, 4)).astype(np.int32)
    coh[::7, 2:] = -1                    # boundary cohesive edges
    coh[1::5, 2] = coh[1::5, 0]          # sides that coincide in UV: not seam in tri-soup mode
    coh[1::5, 3] = coh[1::5, 1]
    coh[2::11, 2] = coh[2::11, 0]        # one side apart, the other together: still seam
    for soup in (False, True):
        for init in (0.0, 0.75):
            got = ctx.seam_sparsity(coh, init, soup)
            want = O.seam_sparsity(coh, V, UV, feat.avgEdgeLen, initSeamLen=init, triSoup=soup)
            print(soup, init, got, want, flush=True)
            assert got == pytest.approx(want, rel=1e-12)
    assert ctx.seam_sparsity(np.zeros((0, 4), np.int32), 0.5) == 0.5


def test_prescreen_and_seam_argument_errors(ctx):
    """Bad ids are rejected before any launch; a short candidate buffer receives the first entries and the
    full count."""
    V, F, UV = height_field_disc(21, seed=0)
    nV = V.shape[0]
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(distort(UV, 0, amp=0.03))
    with pytest.raises(capi.OCError):
        ctx.seam_sparsity(np.array([[0, 1, 2, nV]], np.int32))
    assert ctx.seam_sparsity(np.array([[0, 1, -1, -1]], np.int32), 0.25) == 0.25
    full, ns = ctx.split_candidates(False)
    cand = np.full(3, -7, np.int32)
    n, m = capi.C.c_int(), capi.C.c_int()
    st = ctx.L.oc_split_candidates(ctx.h, 0, 0.6, None, 3, cand, capi.C.byref(n), capi.C.byref(m))
    assert st == capi.OC_OK and n.value == len(full) > 3 and m.value == ns
    np.testing.assert_array_equal(cand, full[:3])


@pytest.mark.parametrize("seed,amp", [(0, 0.0), (1, 0.1), (2, 0.15)])
def test_merge_candidates_match_oracle_bit_exactly(ctx, seed, amp):
    """TriMesh::queryMerge's per-pair screen on random zipper-bottom pairs (a vertex and two of its neighbours,
    both fork columns, plus rows that are not zipper bottoms): status and merged positions bit-identical to the
    oracle, including the pairs projected by the relaxation method and the infeasible ones."""
    from test_oracle_merge import random_rows
    V, F, UV = height_field_disc(41, seed=seed)
    if amp:
        UV = distort(UV, seed, amp=amp)
    rows = random_rows(F, V.shape[0], 3000, seed)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    st, mp, n = ctx.merge_candidates(rows)
    ost, omp, on = O.merge_screen(F, UV, rows, O.RestFeatures(V, F).avgEdgeLen)
    print(seed, amp, "candidates", n, "infeasible", int((st == 2).sum()), flush=True)
    np.testing.assert_array_equal(st, ost)
    assert n == on
    zipper = st != 1
    np.testing.assert_array_equal(mp[zipper], omp[zipper])
    assert (st == 2).any() and (st == 0).any()


def test_merge_candidates_on_a_cut_surface(ctx):
    """The cut icosphere's own cohesive edges at its Tutte map and after Newton steps: the two rows at the ends of
    the cut path are zipper bottoms; the screen equals the oracle's."""
    from optcuts_amd import surfaces as S
    Vc, Fc, cohE, meta = S.workload("icosphere_nu10")
    ctx.set_mesh(Vc, Fc, fixed=(0,))
    UV = ctx.tutte(capi.TUTTE_COMPONENTS, precond=capi.PRECOND_JACOBI, rel_tol=1e-13)[0]
    ael = O.RestFeatures(Vc, Fc).avgEdgeLen
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-12, precond=capi.PRECOND_AMG, pcg_rel_tol=1e-6, pcg_max_iter=2000)
    for k in range(4):
        UVk = ctx.get_uv()
        st, mp, n = ctx.merge_candidates(cohE)
        ost, omp, on = O.merge_screen(Fc, UVk, cohE, ael)
        assert (st != 1).sum() == 2
        np.testing.assert_array_equal(st, ost)
        np.testing.assert_array_equal(mp[st != 1], omp[st != 1])
        ctx.newton_step()
