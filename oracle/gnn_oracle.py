"""CPU oracle of the learned preconditioner (TEST INFRASTRUCTURE ONLY).

Only tests import this module, as the checker of the HIP kernels in
``optcuts_amd/csrc/oc_gnn.hip``. The learned preconditioner has no
counterpart in the reference (OptCuts solves every Newton system with
``EigenLibSolver``'s SimplicialLDLT, ``EigenLibSolver.cpp:71-108``); this file
is the specification of the network, its loss and its gradients, pinned by
central finite differences in float64 (``tests/test_gnn_oracle.py``) and by
the identity ``G = I`` (untrained heads) reducing the preconditioner to
block Jacobi.

Model. For the Newton matrix A (full symmetric 2x2-block CSR over nV
vertices; fixed vertices carry only their diagonal block):
  D_u = A_uu = C_u C_u^T (2x2 Cholesky), A^_uw = C_u^-1 A_uw C_w^-T.
  Node input x_u (4): [1, deg_u / 8, log tr D_u - mean over free vertices,
  fixed_u]. Edge input e_uw (16, for every stored off-diagonal block):
  [A^_uw row-major (4), |A^_uw|_F, w < u, w > u, 0 x 9].
  h = relu(x W_enc^T + b_enc)                                    (32)
  3 layers:  m_uw = relu(bf(h_u | h_w | e_uw) bf(W_m)^T + b_m)     (32)
             a_u  = sum_w m_uw / max(deg_u, 1)
             h_u += relu(bf(h_u | a_u) bf(W_h)^T + b_h)
  heads:     G^_uw = (h_u | h_w | e_uw) W_o^T + b_o  (2x2, lower blocks w < u)
             G^_uu = [[exp d0, 0], [d1, exp d2]],  d = h_u W_d^T + b_d
             (fixed vertices: G^_uu = I)
  preconditioner  M^-1 = C^-T G^T G C^-1  (z = M^-1 r: two sparse products)
  loss  L = sum_j |G^ A^ G^T w_j - w_j|^2 / (k * 2 nV),  w_j Rademacher.
bf() rounds to bfloat16 (round to nearest even) where the GPU feeds MFMA
(the message and update products, forward and backward); with bf16=False
the same function is exact float64, the form the finite differences pin.
"""
import numpy as np
import scipy.sparse as sp

H = 32          # hidden width
NX = 4          # node input features
NE = 16         # edge input features (8 used, zero padded)
KM = 2 * H + NE  # message input width (80)
KH = 2 * H       # update input width (64)
LAYERS = 3

# flat parameter layout (name, shape), in device order
LAYOUT = [("W_enc", (H, NX)), ("b_enc", (H,))]
for _l in range(LAYERS):
    LAYOUT += [("W_m%d" % _l, (H, KM)), ("b_m%d" % _l, (H,)), ("W_h%d" % _l, (H, KH)), ("b_h%d" % _l, (H,))]
LAYOUT += [("W_o", (4, KM)), ("b_o", (4,)), ("W_d", (3, H)), ("b_d", (3,))]
N_PARAMS = int(sum(np.prod(s) for _, s in LAYOUT))


def unflatten(theta):
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def flatten(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def init_params(seed):
    """He-uniform hidden layers, zero heads (G^ = I: block Jacobi)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in LAYOUT:
        if name.startswith("W_") and name not in ("W_o", "W_d"):
            lim = np.sqrt(6.0 / shape[1])
            p[name] = rng.uniform(-lim, lim, size=shape)
        else:
            p[name] = np.zeros(shape)
    # the padded edge-feature columns stay zero
    for l in range(LAYERS):
        p["W_m%d" % l][:, 2 * H + 8:] = 0.0
    return flatten(p).astype(np.float32).astype(np.float64)


def bf16(x):
    """Round to bfloat16 (round to nearest, ties to even) and back to float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """Features and index maps of one Newton system."""

    def __init__(self, rp, ci, vals, fixed):
        rp = np.asarray(rp, np.int64)
        ci = np.asarray(ci, np.int64)
        vals = np.asarray(vals, np.float64).reshape(-1, 2, 2)
        self.nV = nV = len(rp) - 1
        self.fixed = np.asarray(fixed, bool)
        row = np.repeat(np.arange(nV), np.diff(rp))
        diag = row == ci
        dpos = np.full(nV, -1, np.int64)
        dpos[row[diag]] = np.nonzero(diag)[0]
        D = vals[dpos].copy()
        a, b, d = D[:, 0, 0], 0.5 * (D[:, 0, 1] + D[:, 1, 0]), D[:, 1, 1]
        c00 = np.sqrt(a)
        c10 = b / c00
        c11 = np.sqrt(d - c10 * c10)
        self.Cinv = np.zeros((nV, 2, 2))  # C_u^-1 (lower)
        self.Cinv[:, 0, 0] = 1.0 / c00
        self.Cinv[:, 1, 1] = 1.0 / c11
        self.Cinv[:, 1, 0] = -c10 / (c00 * c11)
        off = ~diag
        self.eu = row[off]            # edge e = (u, w), CSR order
        self.ew = ci[off]
        self.E = len(self.eu)
        Ahat = np.einsum("eij,ejk,elk->eil", self.Cinv[self.eu], vals[off], self.Cinv[self.ew])
        self.Ahat = Ahat
        # reverse edge of (u, w) is (w, u)
        key = self.eu * nV + self.ew
        order = np.argsort(key)
        rkey = self.ew * nV + self.eu
        self.rev = order[np.searchsorted(key[order], rkey)]
        self.deg = np.bincount(self.eu, minlength=nV).astype(np.float64)
        trD = a + d
        lt = np.log(trD)
        free = ~self.fixed
        mean = lt[free].mean() if free.any() else 0.0
        self.x = np.stack([np.ones(nV), self.deg / 8.0, np.where(free, lt - mean, 0.0), self.fixed * 1.0], 1)
        ef = np.zeros((self.E, NE))
        ef[:, 0:4] = Ahat.reshape(-1, 4)
        ef[:, 4] = np.sqrt((Ahat ** 2).sum((1, 2)))
        ef[:, 5] = self.ew < self.eu
        ef[:, 6] = self.ew > self.eu
        self.ef = ef
        self.lower = np.nonzero(self.ew < self.eu)[0]
        # scaled matrix A^ as a 2nV x 2nV sparse matrix (identity diagonal blocks)
        self.Ahat_sp = self._block_sparse(self.eu, self.ew, Ahat) + sp.identity(2 * nV, format="csr")

    def _block_sparse(self, r, c, blocks):
        rows = (2 * r[:, None, None] + np.arange(2)[None, :, None]).repeat(2, 2)
        cols = (2 * c[:, None, None] + np.arange(2)[None, None, :]).repeat(2, 1)
        return sp.csr_matrix((blocks.ravel(), (rows.ravel(), cols.ravel())), shape=(2 * self.nV, 2 * self.nV))

    def seg_sum(self, vals_e):
        """sum over each vertex's edges (rows)."""
        out = np.zeros((self.nV,) + vals_e.shape[1:])
        np.add.at(out, self.eu, vals_e)
        return out


def forward(theta, g, bf=True):
    """Returns (G^ lower edge blocks (len(g.lower),2,2), G^ diagonal blocks (nV,2,2), cache)."""
    P = unflatten(theta)
    rnd = bf16 if bf else (lambda a: a)
    relu = lambda a: np.maximum(a, 0.0)
    cache = {"h": []}
    h = relu(g.x @ P["W_enc"].T + P["b_enc"])
    cache["h0"] = h
    deg = np.maximum(g.deg, 1.0)
    for l in range(LAYERS):
        cache["h"].append(h)
        Xe = np.concatenate([h[g.eu], h[g.ew], g.ef], 1)
        m = relu(rnd(Xe) @ rnd(P["W_m%d" % l]).T + P["b_m%d" % l])
        a = g.seg_sum(m) / deg[:, None]
        Xn = np.concatenate([h, a], 1)
        R = relu(rnd(Xn) @ rnd(P["W_h%d" % l]).T + P["b_h%d" % l])
        cache["m%d" % l] = m
        cache["R%d" % l] = R
        h = h + R
    cache["hL"] = h
    L = g.lower
    Xo = np.concatenate([h[g.eu[L]], h[g.ew[L]], g.ef[L]], 1)
    Go = (Xo @ P["W_o"].T + P["b_o"]).reshape(-1, 2, 2)
    d = h @ P["W_d"].T + P["b_d"]
    Gd = np.zeros((g.nV, 2, 2))
    Gd[:, 0, 0] = np.exp(d[:, 0])
    Gd[:, 1, 0] = d[:, 1]
    Gd[:, 1, 1] = np.exp(d[:, 2])
    Gd[g.fixed] = np.eye(2)
    cache["d"] = d
    return Go, Gd, cache


def factor_matrix(g, Go, Gd):
    """G^ as a sparse 2nV x 2nV matrix."""
    L = g.lower
    idx = np.arange(g.nV)
    r = np.concatenate([g.eu[L], idx])
    c = np.concatenate([g.ew[L], idx])
    return g._block_sparse(r, c, np.concatenate([Go, Gd]))


def probes(nV, k, seed):
    """Rademacher probe vectors (k x 2nV), as the GPU draws them (hash of seed, probe, index)."""
    out = np.empty((k, 2 * nV))
    i = np.arange(2 * nV, dtype=np.uint64)
    for j in range(k):
        out[j] = np.where(_hash_bits(i, seed, j) & 1, 1.0, -1.0)
    return out


def _hash_bits(i, seed, j):
    M = np.uint64(0xFFFFFFFF)
    x = (i * np.uint64(0x9E3779B1) + np.uint64(seed) * np.uint64(0x85EBCA77) + np.uint64(j) * np.uint64(0xC2B2AE3D)) & M
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & M
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & M
    x ^= x >> np.uint64(16)
    return x


def loss_and_grad(theta, g, W, bf=True):
    """Loss over the probe vectors W (k x 2nV) and its gradient (flat, like theta)."""
    P = unflatten(theta)
    rnd = bf16 if bf else (lambda a: a)
    Go, Gd, cache = forward(theta, g, bf)
    G = factor_matrix(g, Go, Gd)
    A = g.Ahat_sp
    k, n2 = W.shape
    norm = 1.0 / (k * n2)
    loss = 0.0
    dGo = np.zeros_like(Go)
    dGd = np.zeros_like(Gd)
    L = g.lower
    for w in W:
        y = G.T @ w
        t = A @ y
        e = G @ t - w
        loss += norm * (e @ e)
        gs = 2.0 * norm * e
        q = A @ (G.T @ gs)
        gs2, t2, w2, q2 = gs.reshape(-1, 2), t.reshape(-1, 2), w.reshape(-1, 2), q.reshape(-1, 2)
        u, v = g.eu[L], g.ew[L]
        dGo += gs2[u][:, :, None] * t2[v][:, None, :] + w2[u][:, :, None] * q2[v][:, None, :]
        dGd += gs2[:, :, None] * t2[:, None, :] + w2[:, :, None] * q2[:, None, :]
    grads = {name: np.zeros(shape) for name, shape in LAYOUT}
    d = cache["d"]
    dd = np.stack([dGd[:, 0, 0] * np.exp(d[:, 0]), dGd[:, 1, 0], dGd[:, 1, 1] * np.exp(d[:, 2])], 1)
    dd[g.fixed] = 0.0
    h = cache["hL"]
    grads["W_d"] = dd.T @ h
    grads["b_d"] = dd.sum(0)
    dh = dd @ P["W_d"]
    dGo_f = dGo.reshape(-1, 4)
    Xo = np.concatenate([h[g.eu[L]], h[g.ew[L]], g.ef[L]], 1)
    grads["W_o"] = dGo_f.T @ Xo
    grads["b_o"] = dGo_f.sum(0)
    dXo = dGo_f @ P["W_o"]
    np.add.at(dh, g.eu[L], dXo[:, :H])
    np.add.at(dh, g.ew[L], dXo[:, H:2 * H])
    deg = np.maximum(g.deg, 1.0)
    for l in reversed(range(LAYERS)):
        hin = cache["h"][l]
        R = cache["R%d" % l]
        m = cache["m%d" % l]
        a = g.seg_sum(m) / deg[:, None]
        Xn = np.concatenate([hin, a], 1)
        dZh = dh * (R > 0)
        grads["W_h%d" % l] = rnd(dZh).T @ rnd(Xn)
        grads["b_h%d" % l] = dZh.sum(0)
        dXn = rnd(dZh) @ rnd(P["W_h%d" % l])
        da = dXn[:, H:]
        dh = dh + dXn[:, :H]
        dm = da[g.eu] / deg[g.eu][:, None]
        dZm = dm * (m > 0)
        Xe = np.concatenate([hin[g.eu], hin[g.ew], g.ef], 1)
        grads["W_m%d" % l] = rnd(dZm).T @ rnd(Xe)
        grads["b_m%d" % l] = dZm.sum(0)
        dXe = rnd(dZm) @ rnd(P["W_m%d" % l])
        np.add.at(dh, g.eu, dXe[:, :H])
        np.add.at(dh, g.ew, dXe[:, H:2 * H])
    dZ0 = dh * (cache["h0"] > 0)
    grads["W_enc"] = dZ0.T @ g.x
    grads["b_enc"] = dZ0.sum(0)
    return loss, flatten(grads)


def loss_only(theta, g, W, bf=True):
    Go, Gd, _ = forward(theta, g, bf)
    G = factor_matrix(g, Go, Gd)
    E = (G @ (g.Ahat_sp @ (G.T @ W.T))).T - W
    return float((E * E).sum() / W.size)


def adam(theta, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) as the GPU applies it (t counts from 1)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def precond_apply(g, Go, Gd, r):
    """z = C^-T G^T G C^-1 r."""
    G = factor_matrix(g, Go, Gd)
    r2 = r.reshape(-1, 2)
    s = np.einsum("uij,uj->ui", g.Cinv, r2).ravel()
    y = G.T @ (G @ s)
    return np.einsum("uji,uj->ui", g.Cinv, y.reshape(-1, 2)).ravel()


def pcg_iters(Afull, b, apply_m, tol, maxit=100000):
    """Iterations of preconditioned CG to |r| <= tol |b| (x0 = 0)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p = z.copy()
    rz = r @ z
    bb = b @ b
    for k in range(maxit):
        q = Afull @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if r @ r <= tol * tol * bb:
            return k + 1
        z = apply_m(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return maxit
