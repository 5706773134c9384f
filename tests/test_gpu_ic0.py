"""Block IC(0) preconditioner on the GPU (OC_PRECOND_IC0; the learned form's fallback).

- With sweep counts past the dependency depth, the fixed-point factor and the Jacobi-swept triangular solves
  are exact: M^-1 r equals (L L^T)^-1 r for the sequential IC(0) factor of the tests' helper (pinned in
  tests/test_ic0_helper.py) to 1e-9 relative, on the reference golden 21x21 Newton matrix.
- With the default sweeps M^-1 is symmetric and positive definite (B^T B), so PCG stays valid: checked on
  random vectors to 1e-12.
- IC(0)-PCG solves match the reference SimplicialLDLT goldens to 1e-8 at PCG tolerance 1e-13, in fewer
  iterations than Jacobi-PCG, and a 5k-face Newton trace with OC_PRECOND_IC0 follows the oracle's LDL^T
  trajectory to 1e-9 per iteration.
- The learned form falls back to IC(0)-PCG (pcg_fallback 7) when its own PCG reaches the iteration cap.
"""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd import capi, learned
from optcuts_amd.meshgen import height_field_disc

from _helpers import block_pattern_of_upper, ic0_dense, upper_csr_to_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.ic0_set_sweeps(3, 2)
    c.close()


def _golden_system(ctx, golden, m):
    p = "grid%d_" % m
    V, F, UV = golden[p + "Vrest"], golden[p + "F"], golden[p + "UV"]
    nV = V.shape[0]
    ptr, nbr = O.vertex_neighbors(nV, F)
    ia, ja = ctx.solver_set_pattern(nV, ptr, nbr, fixed=(0,))
    feat = O.RestFeatures(V, F)
    I, J, Vt = O.hessian_triplets(F, O.element_hessians(F, UV, feat), fixed=(0,))
    ctx.solver_update_a(I, J, Vt)
    return ia, ja, ctx.solver_get_a(), p


def test_exact_sweeps_reproduce_sequential_ic0(ctx, golden):
    ia, ja, a, _ = _golden_system(ctx, golden, 21)
    A = upper_csr_to_dense(ia, ja, a)
    L = ic0_dense(A, block_pattern_of_upper(ia, ja))
    ctx.ic0_set_sweeps(3000, 3000)
    try:
        ctx.solver_set_params(capi.PRECOND_IC0, rel_tol=1e-12, max_iter=1000)
        ctx.solver_factorize()
        rng = np.random.default_rng(0)
        for _ in range(3):
            r = rng.normal(size=A.shape[0])
            z = ctx.solver_apply_precond(r)
            zc = np.linalg.solve(L @ L.T, r)
            assert np.linalg.norm(z - zc) <= 1e-9 * np.linalg.norm(zc)
    finally:
        ctx.ic0_set_sweeps(3, 2)


def test_default_sweeps_give_spd_preconditioner(ctx, golden):
    ia, ja, a, _ = _golden_system(ctx, golden, 21)
    ctx.solver_set_params(capi.PRECOND_IC0, rel_tol=1e-12, max_iter=1000)
    ctx.solver_factorize()
    rng = np.random.default_rng(1)
    N = len(ia) - 1
    for _ in range(4):
        r1, r2 = rng.normal(size=N), rng.normal(size=N)
        z1, z2 = ctx.solver_apply_precond(r1), ctx.solver_apply_precond(r2)
        assert abs(r2 @ z1 - r1 @ z2) <= 1e-12 * np.linalg.norm(r1) * np.linalg.norm(z2)
        assert r1 @ z1 > 0.0 and r2 @ z2 > 0.0


@pytest.mark.parametrize("m", [11, 21])
def test_ic0_pcg_matches_reference_golden(ctx, golden, m):
    ia, ja, a, p = _golden_system(ctx, golden, m)
    iters = {}
    for prec in (capi.PRECOND_JACOBI, capi.PRECOND_IC0):
        ctx.solver_set_params(prec, rel_tol=1e-13, max_iter=200000)
        ctx.solver_factorize()
        iters[prec] = 0
        for rhs, sol in ((-golden[p + "g"], golden[p + "d"]), (golden[p + "b2"], golden[p + "d2"])):
            x = ctx.solver_solve(rhs)
            it, rr = ctx.solver_last_stats()
            iters[prec] += it
            assert rr <= 1e-13
            assert np.linalg.norm(x - sol) <= 1e-8 * np.linalg.norm(sol), (m, prec, it)
            assert ctx.solver_last_fallback() == 0
    print("grid", m, "PCG iterations Jacobi", iters[capi.PRECOND_JACOBI], "IC(0)", iters[capi.PRECOND_IC0],
          flush=True)
    assert iters[capi.PRECOND_IC0] < iters[capi.PRECOND_JACOBI]


def test_newton_trace_ic0_matches_oracle_5k_faces(ctx):
    V, F, UV = height_field_disc(51, seed=0)
    nV = V.shape[0]
    rel = 1e-16 * nV / (nV - 1)
    tr, _, _ = O.newton_run(V, F, UV, a0=1.0, max_iter=25, rel_gl2_tol=rel, allow_edec_rel_tol=False)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    ctx.newton_init(a0=1.0, rel_gl2_tol=rel, allow_edec_rel_tol=False, precond=capi.PRECOND_IC0,
                    pcg_rel_tol=1e-12, pcg_max_iter=20000)
    for k in range(1, len(tr)):
        s, done = ctx.newton_step()
        print(k, s["E_total"], tr[k, 0], s["pcg_iters"], flush=True)
        assert s["pcg_fallback"] == 0
        assert s["E_total"] == pytest.approx(tr[k, 0], rel=1e-9), k
        if done:
            break


def test_learned_form_falls_back_to_ic0(ctx):
    """Untrained network (zero output heads: block Jacobi) with an iteration cap that block-Jacobi-PCG misses
    and IC(0)-PCG meets: the step reports pcg_fallback 7 and takes IC(0)-PCG's direction (the energy of a
    plain OC_PRECOND_IC0 step from the same state)."""
    V, F, UV = height_field_disc(41, seed=0)
    ctx.set_mesh(V, F, fixed=(0,))
    tol = 1e-10

    def first_step(precond, cap):
        ctx.set_uv(UV)
        ctx.newton_init(a0=1.0, rel_gl2_tol=1e-12, allow_edec_rel_tol=False, precond=precond, pcg_rel_tol=tol,
                        pcg_max_iter=cap)
        return ctx.newton_step()[0]

    s_bj = first_step(capi.PRECOND_BLOCK_JACOBI, 100000)
    s_ic = first_step(capi.PRECOND_IC0, 100000)
    print("first-step PCG iterations: block Jacobi", s_bj["pcg_iters"], "IC(0)", s_ic["pcg_iters"], flush=True)
    assert s_ic["pcg_fallback"] == 0 and s_ic["pcg_iters"] + 4 < s_bj["pcg_iters"]
    net = capi.Gnn(0)
    try:
        net.set_params(learned.init_params(0))
        ctx.set_learned(net)
        cap = s_ic["pcg_iters"] + 2
        s = first_step(capi.PRECOND_LEARNED, cap)
        print("learned step", s, flush=True)
        assert s["pcg_fallback"] == 7
        assert s["pcg_iters_first"] == cap
        assert s["pcg_iters"] == cap + s_ic["pcg_iters"]
        assert s["E_total"] == pytest.approx(s_ic["E_total"], rel=1e-12)
    finally:
        ctx.set_learned(None)
        net.close()
