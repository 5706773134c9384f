"""The learned-preconditioner oracle pinned without the GPU: its analytic
gradient against central finite differences (float64, no bf16 rounding), the
untrained heads reducing M to block Jacobi, the bf16 rounding rule, and one
Adam step against the textbook update."""
import numpy as np
import pytest

from oracle import gnn_oracle as G
from optcuts_amd.meshgen import height_field_disc
from _helpers import block_system, block_csr_to_sparse, distort


@pytest.fixture(scope="module")
def small_graph():
    V, F, UV = height_field_disc(7, 0)
    UV = distort(UV, 1, 0.1)
    rp, ci, va, fm = block_system(V, F, UV)
    return G.Graph(rp, ci, va, fm), (rp, ci, va, fm)


def _perturbed_params(seed):
    rng = np.random.default_rng(seed)
    P = G.unflatten(G.init_params(seed))
    for name in ("W_o", "b_o", "W_d", "b_d"):
        P[name] = rng.normal(size=P[name].shape) * 0.05
    for name, _ in G.LAYOUT:
        if name.startswith("b_") and name not in ("b_o", "b_d"):
            P[name] = rng.normal(size=P[name].shape) * 0.05
    return G.flatten(P)


def test_layout_size():
    assert G.N_PARAMS == 14599


def test_gradient_matches_finite_differences(small_graph):
    g, _ = small_graph
    th = _perturbed_params(3)
    W = G.probes(g.nV, 3, 7)
    _, grad = G.loss_and_grad(th, g, W, bf=False)
    rng = np.random.default_rng(0)
    live = np.nonzero(np.abs(grad) > 0)[0]
    idx = rng.choice(live, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(th)
        e[i] = h
        fd = (G.loss_only(th + e, g, W, False) - G.loss_only(th - e, g, W, False)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])), (i, fd, grad[i])
    # the loss itself: |G A^ G^T w - w|^2 by dense algebra
    Go, Gd, _ = G.forward(th, g, bf=False)
    Gm = G.factor_matrix(g, Go, Gd).toarray()
    A = g.Ahat_sp.toarray()
    dense = sum(np.sum((Gm @ A @ Gm.T @ w - w) ** 2) for w in W) / W.size
    assert abs(dense - G.loss_only(th, g, W, False)) <= 1e-12 * dense


def test_untrained_heads_give_block_jacobi(small_graph):
    g, (rp, ci, va, fm) = small_graph
    th = G.init_params(0)
    Go, Gd, _ = G.forward(th, g)
    assert np.all(Go == 0.0) and np.allclose(Gd, np.eye(2))
    r = np.random.default_rng(1).normal(size=2 * g.nV)
    z = G.precond_apply(g, Go, Gd, r)
    A = block_csr_to_sparse(rp, ci, va).toarray()
    Dinv = np.zeros_like(A)
    for u in range(g.nV):
        Dinv[2 * u:2 * u + 2, 2 * u:2 * u + 2] = np.linalg.inv(A[2 * u:2 * u + 2, 2 * u:2 * u + 2])
    assert np.allclose(z, Dinv @ r, rtol=1e-12, atol=1e-12)
    # A^ has identity diagonal blocks and is C^-1 A C^-T
    C = np.zeros_like(A)
    for u in range(g.nV):
        C[2 * u:2 * u + 2, 2 * u:2 * u + 2] = np.linalg.cholesky(A[2 * u:2 * u + 2, 2 * u:2 * u + 2])
    Ci = np.linalg.inv(C)
    assert np.allclose(g.Ahat_sp.toarray(), Ci @ A @ Ci.T, atol=1e-10)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159, 0.0])
    y = G.bf16(x)
    # bf16 keeps 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 + 2 ** -6 and y[3] == 1.0
    assert y[4] == -3.140625 and y[5] == 0.0


def test_adam_step_textbook():
    rng = np.random.default_rng(2)
    th, g = rng.normal(size=10), rng.normal(size=10)
    m, v = np.zeros(10), np.zeros(10)
    th1, m1, v1 = G.adam(th, g, m, v, 1, 1e-3)
    # first step: m_hat = g, v_hat = g^2 -> step = lr * sign(g) (up to eps)
    assert np.allclose(th1, th - 1e-3 * g / (np.abs(g) + 1e-8))


def test_probe_hash_is_rademacher():
    W = G.probes(5000, 2, 11)
    assert set(np.unique(W)) == {-1.0, 1.0}
    assert abs(W.mean()) < 0.03 and abs((W[0] * W[1]).mean()) < 0.03
