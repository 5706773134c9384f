"""GPU parity of the learned preconditioner against its CPU oracle
(oracle/gnn_oracle.py, with the same bf16 rounding of the matrix-core
operands): factor, preconditioner application, loss, gradients and Adam; then
the Newton driver with the learned preconditioner against the reference
trajectory, and what training gains over the untrained (block-Jacobi)
network.

Tolerances: the GPU accumulates in FP32 (MFMA and vector sums) where the
oracle uses float64, and both round the same operands to bf16, so factor and
loss agree to ~1e-5 relative and whole gradient tensors to 2e-3 in norm (a
ReLU or bf16 rounding can flip on an input within FP32 rounding of a
boundary)."""
import numpy as np
import pytest

from oracle import gnn_oracle as G
from oracle import oracle as O
from optcuts_amd import capi
from optcuts_amd import learned
from optcuts_amd.meshgen import height_field_disc
from _helpers import block_system, block_csr_to_sparse, distort

pytestmark = pytest.mark.gpu


def _params(seed, head_scale=0.05):
    rng = np.random.default_rng(seed)
    P = G.unflatten(learned.init_params(seed).astype(np.float64))
    for name in ("W_o", "b_o", "W_d", "b_d"):
        P[name] = rng.normal(size=P[name].shape) * head_scale
    for name, _ in G.LAYOUT:
        if name.startswith("b_") and name not in ("b_o", "b_d"):
            P[name] = rng.normal(size=P[name].shape) * 0.05
    return G.flatten(P).astype(np.float32)


@pytest.fixture(scope="module", params=[9, 31])
def system(request):
    m = request.param
    V, F, UV = height_field_disc(m, 0)
    UV = distort(UV, 2, 0.1)
    rp, ci, va, fm = block_system(V, F, UV)
    print("gnn system grid", m, "vertices", len(rp) - 1, "blocks", len(ci), flush=True)
    return rp, ci, va, fm


@pytest.fixture(scope="module")
def net():
    n = capi.Gnn(0)
    yield n
    n.close()


def test_param_layout(net):
    assert net.n_params == G.N_PARAMS
    th = _params(1)
    net.set_params(th)
    assert np.array_equal(net.get_params(), th)


def test_factor_matches_oracle(net, system):
    rp, ci, va, fm = system
    th = _params(4)
    net.set_params(th)
    net.set_system(rp, ci, va, fm)
    Gb = net.forward()
    g = G.Graph(rp, ci, va, fm)
    Go, Gd, _ = G.forward(th.astype(np.float64), g, bf=True)
    # oracle blocks per CSR slot
    ref = np.zeros((len(ci), 2, 2))
    off = np.nonzero(np.repeat(np.arange(len(rp) - 1), np.diff(rp)) != ci)[0]
    low = off[g.lower]
    ref[low] = Go
    diag = np.nonzero(np.repeat(np.arange(len(rp) - 1), np.diff(rp)) == ci)[0]
    ref[diag] = Gd
    scale = np.abs(ref).max()
    err = np.abs(Gb - ref).max()
    print("factor max err", err, "scale", scale, flush=True)
    assert err <= 2e-4 * scale


def test_precond_apply_matches_oracle(net, system):
    rp, ci, va, fm = system
    th = _params(5)
    net.set_params(th)
    net.set_system(rp, ci, va, fm)
    net.forward()
    g = G.Graph(rp, ci, va, fm)
    Go, Gd, _ = G.forward(th.astype(np.float64), g, bf=True)
    r = np.random.default_rng(3).normal(size=2 * g.nV)
    z = net.precond_apply(r)
    zr = G.precond_apply(g, Go, Gd, r)
    assert np.linalg.norm(z - zr) <= 1e-4 * np.linalg.norm(zr)
    # symmetric positive definite operator: r1.M r2 == r2.M r1, r.M r > 0
    r2 = np.random.default_rng(4).normal(size=2 * g.nV)
    z2 = net.precond_apply(r2)
    assert abs(r @ z2 - r2 @ z) <= 1e-10 * np.sqrt((r @ z) * (r2 @ z2))
    assert r @ z > 0


def test_loss_and_gradient_match_oracle(net, system):
    rp, ci, va, fm = system
    th = _params(6)
    net.set_params(th)
    net.set_system(rp, ci, va, fm)
    net.zero_grad()
    k, seed = 3, 12345
    loss = net.loss_grad(k, seed)
    grad = net.get_grads()
    g = G.Graph(rp, ci, va, fm)
    W = G.probes(g.nV, k, seed)
    lr, gr = G.loss_and_grad(th.astype(np.float64), g, W, bf=True)
    print("loss gpu", loss, "oracle", lr, flush=True)
    assert abs(loss - lr) <= 1e-4 * lr
    P, Pr = G.unflatten(grad.astype(np.float64)), G.unflatten(gr)
    for name, _ in G.LAYOUT:
        a, b = P[name], Pr[name]
        nb = np.linalg.norm(b)
        err = np.linalg.norm(a - b)
        print("grad", name, "rel err %.2e" % (err / max(nb, 1e-30)), "norm %.3e" % nb, flush=True)
        assert err <= 2e-3 * nb + 1e-7 * np.linalg.norm(gr), name
    # accumulation: a second call adds the same gradient
    net.loss_grad(k, seed)
    assert np.allclose(net.get_grads(), 2 * grad, rtol=1e-5, atol=1e-6 * np.abs(grad).max())


def test_adam_matches_oracle(system):
    rp, ci, va, fm = system
    th = _params(7)
    net = capi.Gnn(0)  # fresh optimizer state
    request_close = net.close
    net.set_params(th)
    net.set_system(rp, ci, va, fm)
    net.zero_grad()
    net.loss_grad(2, 99)
    grad = net.get_grads().astype(np.float64)
    net.adam_step(1e-3, 1)
    th1, _, _ = G.adam(th.astype(np.float64), grad, np.zeros_like(grad), np.zeros_like(grad), 1, 1e-3)
    try:
        assert np.abs(net.get_params() - th1).max() <= 1e-6
    finally:
        request_close()


def test_training_beats_untrained_network():
    """Train on Hessians from the method's own Newton iterations on small
    discs; on held-out systems (another seed and size) the trained factor
    needs fewer PCG iterations than the untrained one (block Jacobi)."""
    train_sys = learned.collect_systems(learned.disc_meshes([21, 31], [1, 2]), iters=6)
    test_sys = learned.collect_systems(learned.disc_meshes([41], [5]), iters=6)
    net = capi.Gnn(0)
    try:
        th0 = learned.init_params(0)
        net.set_params(th0)
        rng = np.random.default_rng(0)
        rhs = [rng.normal(size=2 * (len(s[0]) - 1)) for s in test_sys]
        base = [learned.pcg_iterations(net, s, b) for s, b in zip(test_sys, rhs)]
        hist = learned.train(net, train_sys, epochs=15, lr=2e-3, probes=4)
        trained = [learned.pcg_iterations(net, s, b) for s, b in zip(test_sys, rhs)]
        print("train loss first/last", hist[0], hist[-1], flush=True)
        print("PCG iterations untrained", base, "trained", trained, flush=True)
        assert hist[-1] < hist[0]
        assert sum(trained) < 0.9 * sum(base)
    finally:
        net.close()


def test_training_on_closed_surface_systems():
    """Training on the scaffolded joint-mode systems of cut closed surfaces (mesh + air rows, genus 0 and 3) and
    evaluating on systems of other resolutions. Measured: the loss falls by three orders of magnitude, yet the
    trained factor needs more PCG iterations than block Jacobi on the held-out icosphere systems (482 -> 1051 on the
    first), and on the genus-3 systems of the first Tutte steps neither reaches 1e-6 in 5,000 iterations: the
    Frobenius objective does not track CG convergence on these spectra. The test keeps what must hold: the loss
    decreases and the trained factor is still a usable (positive definite) preconditioner where block Jacobi
    converges."""
    train_sys = learned.collect_closed_systems(["icosphere_nu12", "tori_h0.12"], iters=8)
    test_sys = learned.collect_closed_systems(["icosphere_nu16"], iters=4)
    net = capi.Gnn(0)
    try:
        net.set_params(learned.init_params(0))
        rng = np.random.default_rng(1)
        rhs = [rng.normal(size=2 * (len(s[0]) - 1)) for s in test_sys]
        base = [learned.pcg_iterations(net, s, b, tol=1e-6, maxit=5000) for s, b in zip(test_sys, rhs)]
        hist = learned.train(net, train_sys, epochs=15, lr=2e-3, probes=4)
        trained = [learned.pcg_iterations(net, s, b, tol=1e-6, maxit=5000) for s, b in zip(test_sys, rhs)]
        print("closed: train loss first/last", hist[0], hist[-1], flush=True)
        print("closed: PCG iterations untrained", base, "trained", trained, flush=True)
        assert hist[-1] < 0.1 * hist[0]
        assert max(base) < 5000 and max(trained) < 5000
    finally:
        net.close()


def test_newton_learned_matches_reference_trace():
    """The learned preconditioner changes only how PCG converges: to a tight
    tolerance the Newton trajectory matches the oracle's LDL^T run to 1e-9."""
    V, F, UV = height_field_disc(21, 0)
    tr, _, _ = O.newton_run(V, F, UV, a0=1.0, max_iter=5, rel_gl2_tol=1e-30, allow_edec_rel_tol=False)
    ctx = capi.Context(0)
    net = capi.Gnn(0)
    try:
        net.set_params(_params(8, head_scale=0.02))
        ctx.set_mesh(V, F, fixed=(0,))
        ctx.set_uv(UV)
        ctx.set_learned(net)
        E0 = ctx.newton_init(a0=1.0, rel_gl2_tol=1e-30, allow_edec_rel_tol=False, precond=capi.PRECOND_LEARNED,
                             pcg_rel_tol=1e-12, pcg_max_iter=100000)
        assert abs(E0 - tr[0, 0]) <= 1e-12 * tr[0, 0]
        for k in range(1, len(tr)):
            s, _ = ctx.newton_step()
            print("learned iter", k, s["E_total"], "pcg", s["pcg_iters"], "fallback", s["pcg_fallback"], flush=True)
            assert abs(s["E_total"] - tr[k, 0]) <= 1e-9 * tr[k, 0]
    finally:
        ctx.close()
        net.close()


@pytest.mark.parametrize("precond,tol", [("learned", 1e-4), ("learned", 5e-3)])
def test_mixed_form_200k_reaches_target_without_jacobi(precond, tol):
    """Learned factor in the mixed form at 200k faces, run to the workload's E_SD target (1e-6): a solve that
    breaks down on the FP32-assembled matrix is retried on the matrix rebuilt from FP64 element terms with a
    newly predicted factor (pcg_fallback 4, after which the solve's Newton systems stay FP64), so no step falls
    back to Jacobi-PCG or truncates. 5e-3 is the benchmark's PCG tolerance for this form (inexact Newton: about
    twice the Newton iterations of 1e-4, each much cheaper)."""
    import json
    import os
    e_star = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workload_targets.json")))["E_SD_star"]
    V, F, UV = height_field_disc(317, 0)
    nV = V.shape[0]
    ctx = capi.Context(0)
    net = None
    try:
        if precond == "learned":
            net = capi.Gnn(0)
            net.set_params(learned.load_weights())
            ctx.set_learned(net)
        ctx.set_mesh(V, F, fixed=(0,))
        ctx.set_uv(UV)
        ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16 * nV / (nV - 1), allow_edec_rel_tol=False,
                        precond=capi.PRECOND_LEARNED if precond == "learned" else capi.PRECOND_AMG,
                        pcg_rel_tol=tol, pcg_max_iter=200000 if precond == "learned" else 1000,
                        precision=capi.PRECISION_MIXED)
        codes = []
        for k in range(1, 61):
            s, done = ctx.newton_step()
            codes.append(s["pcg_fallback"])
            print(precond, "mixed iter", k, s["E_SD"], "pcg", s["pcg_iters"], "fallback", s["pcg_fallback"],
                  flush=True)
            if abs(s["E_SD"] - e_star) <= 1e-6 * e_star or done:
                break
        assert abs(ctx.energy() - e_star) <= 1e-6 * e_star
        assert set(codes) <= {0, 4}, codes
    finally:
        ctx.close()
        if net is not None:
            net.close()
