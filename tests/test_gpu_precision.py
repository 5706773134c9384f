"""GPU checks of the mixed-precision element path (FP32 element arithmetic) against the FP64 oracle, and of the
reference's edge cases on the device.

FP32 element error model. The element is evaluated in FP32 on the vertex positions relative to its first vertex
(the differences are formed in FP64 and rounded once). The energy's sensitive quantity is the UV area
a = (ax by - ay bx) / 2, whose relative rounding error is about eps32 * kappa_t with
kappa_t = (|ax by| + |ay bx|) / |ax by - ay bx| (1 for a right isoceles triangle, large when the triangle is
nearly flat). Energy, gradient and Hessian scale as E_t, E_t / h_t and E_t / h_t^2 with h_t = 2|a| / l_max the
smallest height of the UV triangle. The test bounds, per element,
    |E32 - E64| / E64                 <= C_E eps32 (kappa_t + 1)
    max|G32 - G64| / (E64 / h_t)      <= C_G eps32 (kappa_t + 1)
    max|H32 - H64| / (E64 / h_t^2)    <= C_H eps32 (kappa_t + 1)
with the constants in BOUNDS, on distorted maps that include triangles brought to within 1e-4 of inverting
(kappa_t up to ~1e5). The PSD projection is 1-Lipschitz in the Frobenius norm, so projecting both Hessians adds
no amplification beyond the FP32 eigen-solve's own rounding.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd import capi
from optcuts_amd.meshgen import band_air_mesh, boundary_loop_ccw, height_field_disc

from _helpers import distort, grid_air_mesh
from test_scaffold_oracle import TWO_TRI_IJ, check_two_triangle_triplets, two_triangle_case

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
BOUNDS = {"E": 16.0, "G": 16.0, "H": 32.0}


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def near_degenerate_maps(m=31, seed=5):
    """A distorted disc map, and maps pushed along random directions to 0.9, 0.99 and 0.9999 of the first
    inversion (SymDirichletEnergy::initStepSize's bound), so some triangles are nearly flat."""
    V, F, UV = height_field_disc(m, seed=seed)
    UV = distort(UV, seed, amp=0.15)
    out = [("distorted", UV)]
    rng = np.random.default_rng(seed)
    for f in (0.9, 0.99, 0.9999):
        d = rng.normal(size=2 * V.shape[0]) * 0.01
        d[:2] = 0.0
        a = O.init_step_size(F, UV, d, 1e30)
        UVd = UV + f * a * d.reshape(-1, 2)
        assert O.check_inversion(F, UVd)
        out.append(("to %.4f of inversion" % f, UVd))
    return V, F, out


def kappa_and_height(F, UV):
    P = UV[F]
    a = P[:, 1] - P[:, 0]
    b = P[:, 2] - P[:, 0]
    cr = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    kappa = (np.abs(a[:, 0] * b[:, 1]) + np.abs(a[:, 1] * b[:, 0])) / np.abs(cr)
    lmax = np.max(np.stack([np.linalg.norm(a, axis=1), np.linalg.norm(b, axis=1),
                            np.linalg.norm(P[:, 2] - P[:, 1], axis=1)]), axis=0)
    return kappa, np.abs(cr) / lmax


def test_fp64_element_terms_match_oracle(ctx):
    V, F, maps = near_degenerate_maps(m=21, seed=2)
    feat = O.RestFeatures(V, F)
    ctx.set_mesh(V, F, fixed=(0,))
    eps64 = float(np.finfo(np.float64).eps)
    for name, UV in maps:
        ctx.set_uv(UV)
        E, G, H = ctx.element_terms(capi.PRECISION_FP64)
        E0 = O.energy_per_elem(F, UV, feat)
        G0 = O.local_gradients(F, UV, feat).reshape(-1, 6)
        H0 = O.element_hessians(F, UV, feat)
        # the same arithmetic up to FMA contraction: the error model of the FP32 test with eps64; the projected
        # Hessians come from two different eigensolvers (the device's 4x4 reduced Jacobi, the oracle's 6x6 one),
        # whose FP64 roundings differ by up to ~60 such units (measured 57)
        kappa, h = kappa_and_height(F, UV)
        unit = eps64 * (kappa + 1.0)
        assert (np.abs(E - E0) / E0 / unit).max() <= BOUNDS["E"], name
        assert (np.abs(G - G0).max(axis=1) / (E0 / h) / unit).max() <= BOUNDS["G"], name
        assert (np.abs(H - H0).max(axis=(1, 2)) / (E0 / h ** 2) / unit).max() <= 256.0, name
        # and on the ordinary (distorted) map simply to 1e-12
        if name == "distorted":
            np.testing.assert_allclose(E, E0, rtol=1e-12)


def test_fp32_element_terms_bounded_against_fp64_oracle(ctx):
    V, F, maps = near_degenerate_maps()
    feat = O.RestFeatures(V, F)
    ctx.set_mesh(V, F, fixed=(0,))
    worst = {"E": 0.0, "G": 0.0, "H": 0.0}
    for name, UV in maps:
        ctx.set_uv(UV)
        E, G, H = ctx.element_terms(capi.PRECISION_MIXED)
        E0 = O.energy_per_elem(F, UV, feat)
        G0 = O.local_gradients(F, UV, feat).reshape(-1, 6)
        H0 = O.element_hessians(F, UV, feat)
        kappa, h = kappa_and_height(F, UV)
        unit = EPS32 * (kappa + 1.0)
        rE = np.abs(E - E0) / E0 / unit
        rG = np.abs(G - G0).max(axis=1) / (E0 / h) / unit
        rH = np.abs(H - H0).max(axis=(1, 2)) / (E0 / h ** 2) / unit
        # the FP32 path is not the FP64 path in disguise
        assert np.abs(E - E0).max() > 0.0
        for k, r in (("E", rE), ("G", rG), ("H", rH)):
            worst[k] = max(worst[k], float(r.max()))
        print(name, "kappa max %.3g" % kappa.max(), "ratios E %.3g G %.3g H %.3g" % (rE.max(), rG.max(), rH.max()),
              "max rel E err %.3g" % (np.abs(E - E0) / E0).max(), flush=True)
    print("worst normalised errors", worst, flush=True)
    for k in worst:
        assert worst[k] <= BOUNDS[k], (k, worst[k])


def test_fp32_element_terms_independent_of_uv_offset(ctx):
    """Relative coordinates: moving the whole map by a large offset changes the FP32 terms only by the rounding
    of the FP64 differences (an element evaluated on absolute FP32 coordinates would lose ~1e-4 here)."""
    V, F, maps = near_degenerate_maps(m=21, seed=3)
    _, UV = maps[1]
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    E, G, H = ctx.element_terms(capi.PRECISION_MIXED)
    ctx.set_uv(UV + np.array([1000.0, -700.0]))
    E2, G2, H2 = ctx.element_terms(capi.PRECISION_MIXED)
    assert np.abs(E2 - E).max() <= 1e-6 * np.abs(E).max()
    assert np.abs(G2 - G).max() <= 1e-6 * np.abs(G).max()
    assert np.abs(H2 - H).max() <= 1e-6 * np.abs(H).max()


def test_mixed_amg_newton_trace_matches_oracle_5k_faces(ctx):
    """The mixed form (FP32 element terms, FP32-stored multigrid hierarchy, FP64 PCG) at a tight PCG tolerance
    follows the oracle's FP64 LDL^T trajectory on the 5,000-face disc to 1e-5 relative in energy per iteration
    (the FP32 element terms perturb each Newton direction by ~1e-6 relative; measured worst 2.8e-6) and ends at
    the oracle's optimum to 1e-7."""
    V, F, UV = height_field_disc(51, seed=0)
    nV = V.shape[0]
    rel = 1e-16 * nV / (nV - 1)
    tr, _, _ = O.newton_run(V, F, UV, a0=1.0, max_iter=25, rel_gl2_tol=rel, allow_edec_rel_tol=False)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    E0 = ctx.newton_init(a0=1.0, rel_gl2_tol=rel, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG,
                         pcg_rel_tol=1e-12, pcg_max_iter=4000, precision=capi.PRECISION_MIXED)
    assert E0 == pytest.approx(tr[0, 0], rel=2e-6)
    worst = 0.0
    s = None
    for k in range(1, len(tr)):
        s, done = ctx.newton_step()
        err = abs(s["E_total"] - tr[k, 0]) / tr[k, 0]
        worst = max(worst, err)
        print(k, s["E_total"], tr[k, 0], "rel %.2e" % err, "pcg", s["pcg_iters"], "fallback", s["pcg_fallback"],
              flush=True)
        assert err <= 1e-5, k
        if done:
            break
    print("worst per-iteration relative energy error %.3g" % worst, flush=True)
    assert abs(s["E_total"] - tr[-1, 0]) <= 1e-7 * tr[-1, 0]
    UVg = ctx.get_uv()
    assert O.check_inversion(F, UVg)
    # the reached map re-evaluated in FP64
    assert O.energy(F, UVg, O.RestFeatures(V, F)) == pytest.approx(tr[-1, 0], rel=1e-7)


def test_triplet_order_two_triangles_by_hand(ctx):
    V, F, UV = two_triangle_case()
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    I, J, Vv = ctx.hessian_triplets()
    H = O.element_hessians(F, UV, O.RestFeatures(V, F))
    assert list(zip(I.tolist(), J.tolist())) == TWO_TRI_IJ
    # the device values equal the oracle's triplet values to FP64 rounding (FMA contraction), and the oracle's
    # values sit at the hand-derived positions
    Io, Jo, Vo = O.hessian_triplets(F, H, fixed=(0,))
    assert np.abs(Vv - Vo).max() <= 1e-12 * np.abs(Vo).max()
    check_two_triangle_triplets(Io, Jo, Vo, H)


def test_zero_area_rest_triangle_is_rejected(ctx):
    """TriMesh::computeFeatures exits on a zero-area rest triangle (TriMesh.cpp:368-372): OC_ERR_DEGENERATE."""
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0], [1.0, 1.0, 0.0]])
    F = np.array([[0, 1, 3], [1, 2, 3], [0, 2, 1]], np.int32)
    with pytest.raises(capi.OCError) as ei:
        ctx.set_mesh(V, F, fixed=(0,))
    assert ei.value.status == -6


def test_clockwise_initial_map_is_rejected(ctx):
    """Optimizer's constructor exits on an inverted initial map (Optimizer.cpp:65-67): OC_ERR_INVERTED."""
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    F = np.array([[0, 1, 2]], np.int32)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(np.array([[0.0, 0.0], [0.0, 1.0], [1.0, 0.0]]))
    with pytest.raises(capi.OCError) as ei:
        ctx.newton_init(a0=1.0)
    assert ei.value.status == -5


def test_one_triangle_newton_reaches_isometry(ctx):
    """One triangle with only vertex 0 fixed: the energy is invariant under rotations about vertex 0, so the
    projected Newton matrix can be singular along that rotation (a direct LDL^T, the oracle's and the
    reference's, then meets a zero pivot); Jacobi-PCG finds a direction in its range. The loop decreases the
    energy monotonically to the isometry, E_SD = 4, with the map counter-clockwise throughout."""
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.3], [0.2, 0.9, -0.1]])
    F = np.array([[0, 1, 2]], np.int32)
    UV = np.array([[0.0, 0.0], [0.5, 0.1], [0.9, 1.7]])
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    E = ctx.newton_init(a0=1.0, rel_gl2_tol=1e-12, allow_edec_rel_tol=True, pcg_rel_tol=1e-14, pcg_max_iter=1000)
    assert E == pytest.approx(O.energy(F, UV, O.RestFeatures(V, F)), rel=1e-13)
    for k in range(40):
        s, done = ctx.newton_step()
        if s["converged"] or done:
            break
        assert s["E_total"] <= E
        E = s["E_total"]
        assert O.check_inversion(F, ctx.get_uv())
    assert ctx.energy() == pytest.approx(4.0, rel=1e-8)


def test_two_triangles_newton_matches_oracle(ctx):
    """Two triangles through the device Newton loop: the oracle's trajectory per iteration, ending at the
    isometry (E_SD = 4)."""
    case = "two_triangles"
    V, F, UV = two_triangle_case()
    tr, UVo, _ = O.newton_run(V, F, UV, a0=1.0, max_iter=30, rel_gl2_tol=1e-12, allow_edec_rel_tol=True)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    E0 = ctx.newton_init(a0=1.0, rel_gl2_tol=1e-12, allow_edec_rel_tol=True, pcg_rel_tol=1e-14,
                         pcg_max_iter=1000)
    assert E0 == pytest.approx(tr[0, 0], rel=1e-13)
    for k in range(1, len(tr)):
        s, done = ctx.newton_step()
        print(case, k, s["E_total"], tr[k, 0], flush=True)
        if s["converged"]:
            break
        assert s["E_total"] == pytest.approx(tr[k, 0], rel=1e-9), k
        if done:
            break
    assert ctx.energy() == pytest.approx(4.0, rel=1e-6)
    np.testing.assert_allclose(ctx.get_uv(), UVo, rtol=0, atol=1e-6)


def test_fixed_vertex_block_with_scaffold(ctx):
    """The device Newton system's fixed-vertex block is (a0 + w_scaf / F_air) I when the fixed vertex is also an
    air-mesh vertex, and identity-scaled a0 I for a fixed vertex away from the air mesh."""
    m = 11
    V, F, UV = height_field_disc(m, seed=3)
    airUV, Fair, bnd = grid_air_mesh(m, rings=2)
    a0, w = 0.3, 0.02
    centre = (m // 2) * m + m // 2
    ctx.set_mesh(V, F, fixed=(0, centre))
    ctx.set_uv(UV)
    ctx.set_air_mesh(airUV, Fair, bnd, 0.0, w)
    ctx.newton_init(a0=a0, rel_gl2_tol=1e-30, allow_edec_rel_tol=False, pcg_rel_tol=1e-10, pcg_max_iter=10000)
    ctx.newton_step()
    rp, ci, vals, b, fm = ctx.get_system()
    for v, d in ((0, a0 + w / Fair.shape[0]), (centre, a0)):
        assert fm[v] == 1 and rp[v + 1] - rp[v] == 1 and ci[rp[v]] == v
        np.testing.assert_array_equal(vals[rp[v]], d * np.eye(2))


def test_rebuild_keeps_trace_and_writes_reference_line_count(ctx):
    """A scaffolded solve that replaces the air mesh after every step (Optimizer.cpp:251-257) through
    oc_newton_rebuild writes the energy trace as the reference does: one line at the start (precompute,
    Optimizer.cpp:197-198) and one per step that did not stop (:645-648), one gradient line per step."""
    m = 21
    V, F, UV = height_field_disc(m, seed=1)
    loop = boundary_loop_ccw(F)
    with tempfile.TemporaryDirectory() as d:
        ctx.set_mesh(V, F, fixed=(0,))
        ctx.set_uv(UV)
        ctx.trace_open(d)
        ctx.set_air_mesh(*band_air_mesh(UV, loop, rings=2, seed=0), 0.0, 0.01)
        ctx.newton_init(a0=1.0, rel_gl2_tol=1e-30, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG,
                        pcg_rel_tol=1e-6, pcg_max_iter=1000)
        n_steps = 4
        for k in range(n_steps):
            s, done = ctx.newton_step()
            assert not done
            # a new air mesh at the current UV (fresh diagonals)
            ctx.set_air_mesh(*band_air_mesh(ctx.get_uv(), loop, rings=2, seed=k + 1), 0.0, 0.01)
            E = ctx.newton_rebuild()
            assert E > 0
        ctx.trace_close()
        eLines = open(os.path.join(d, "energyValPerIter.txt")).read().splitlines()
        gLines = open(os.path.join(d, "gradientPerIter.txt")).read().splitlines()
    assert len(eLines) == 1 + n_steps
    assert len(gLines) == n_steps
