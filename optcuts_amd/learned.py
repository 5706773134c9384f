"""Training and use of the learned Newton-system preconditioner.

The network, its loss and its backward pass run on the GPU inside the C ABI
(``oc_gnn_*``); this module is the host-side driver: parameter
initialisation, collection of training systems from the method's own Newton
iterations, the Adam loop, evaluation by PCG iteration counts, and weight
files.

    python -m optcuts_amd.learned --out optcuts_amd/weights/gnn_sd.npy
"""
import argparse
import os
import time

import numpy as np

from . import capi
from .meshgen import height_field_disc

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "gnn_sd.npy")

H, NX, NE, LAYERS = 32, 4, 16, 3
KM, KH = 2 * H + NE, 2 * H
LAYOUT = [("W_enc", (H, NX)), ("b_enc", (H,))]
for _l in range(LAYERS):
    LAYOUT += [("W_m%d" % _l, (H, KM)), ("b_m%d" % _l, (H,)), ("W_h%d" % _l, (H, KH)), ("b_h%d" % _l, (H,))]
LAYOUT += [("W_o", (4, KM)), ("b_o", (4,)), ("W_d", (3, H)), ("b_d", (3,))]


def init_params(seed=0):
    """He-uniform hidden layers; zero output heads, so the untrained network
    is exactly the block-Jacobi preconditioner. Padded edge-feature columns
    of W_m (inputs 72..79) are zero and stay zero (their gradient vanishes)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in LAYOUT:
        if name.startswith("W_") and name not in ("W_o", "W_d"):
            lim = np.sqrt(6.0 / shape[1])
            w = rng.uniform(-lim, lim, size=shape)
            if name.startswith("W_m"):
                w[:, 2 * H + 8:] = 0.0
        else:
            w = np.zeros(shape)
        parts.append(w.ravel())
    return np.concatenate(parts).astype(np.float32)


def collect_systems(meshes, iters, precond=capi.PRECOND_AMG, pcg_tol=1e-4, device=0):
    """Newton systems (rowPtr, colIdx, vals, fixedMask) from `iters` Newton
    iterations of each (V, F, UV) mesh (vertex 0 fixed, mesh-only objective)."""
    out = []
    ctx = capi.Context(device)
    try:
        for V, F, UV in meshes:
            nV = V.shape[0]
            ctx.set_mesh(V, F, fixed=(0,))
            ctx.set_uv(UV)
            ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16 * nV / (nV - 1), allow_edec_rel_tol=False, precond=precond,
                            pcg_rel_tol=pcg_tol, pcg_max_iter=1000 if precond == capi.PRECOND_AMG else 200000)
            for _ in range(iters):
                s, done = ctx.newton_step()
                if s["converged"]:
                    break
                rp, ci, va, b, fm = ctx.get_system()
                out.append((rp, ci, va.copy(), fm))
                if done:
                    break
    finally:
        ctx.close()
    return out


def disc_meshes(grids, seeds):
    return [height_field_disc(m, seed=s) for m in grids for s in seeds]


def train(net, systems, epochs, lr=1e-3, probes=4, seed=0, log=None):
    """Adam over the systems (one step per system, shuffled each epoch)."""
    rng = np.random.default_rng(seed)
    t = 0
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(systems))
        tot = 0.0
        for i in order:
            rp, ci, va, fm = systems[i]
            net.set_system(rp, ci, va, fm)
            net.zero_grad()
            tot += net.loss_grad(probes, seed=int(rng.integers(1 << 31)))
            t += 1
            net.adam_step(lr, t)
        hist.append(tot / len(systems))
        if log:
            log("epoch %d mean loss %.6g" % (ep, hist[-1]))
    return hist


def pcg_iterations(net, system, rhs, tol=1e-6, maxit=20000):
    """PCG iterations with the network's preconditioner on one system (host
    loop; an evaluation tool, not the solver the Newton driver uses)."""
    import scipy.sparse as sp
    rp, ci, va, fm = system
    nV = len(rp) - 1
    row = np.repeat(np.arange(nV), np.diff(rp))
    r_ = (2 * row[:, None, None] + np.arange(2)[None, :, None]).repeat(2, 2)
    c_ = (2 * ci[:, None, None] + np.arange(2)[None, None, :]).repeat(2, 1)
    A = sp.csr_matrix((np.asarray(va).ravel(), (r_.ravel(), c_.ravel())), shape=(2 * nV, 2 * nV))
    net.set_system(rp, ci, va, fm)
    net.forward()
    x = np.zeros(2 * nV)
    r = rhs.copy()
    z = net.precond_apply(r)
    p = z.copy()
    rz = r @ z
    bb = rhs @ rhs
    for k in range(maxit):
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        if r @ r <= tol * tol * bb:
            return k + 1
        z = net.precond_apply(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return maxit


def save(theta, path=DEFAULT_WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(theta, np.float32))


def load_weights(path=DEFAULT_WEIGHTS):
    return np.load(path, allow_pickle=False).astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--grids", default="31,41,51,71")
    ap.add_argument("--seeds", default="1,2,3")
    ap.add_argument("--iters", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--probes", type=int, default=4)
    args = ap.parse_args()
    t0 = time.time()
    meshes = disc_meshes([int(x) for x in args.grids.split(",")], [int(x) for x in args.seeds.split(",")])
    systems = collect_systems(meshes, args.iters)
    print("collected %d systems in %.1f s" % (len(systems), time.time() - t0), flush=True)
    net = capi.Gnn(0)
    net.set_params(init_params(0))
    train(net, systems, args.epochs, lr=args.lr, probes=args.probes, log=lambda m: print(m, flush=True))
    save(net.get_params(), args.out)
    print("saved %s (%.1f s)" % (args.out, time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
