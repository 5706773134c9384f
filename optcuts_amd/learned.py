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


def collect_closed_systems(names, iters, device=0, pcg_tol=1e-3, pcg_max_iter=100):
    """Newton systems (mesh + air rows) of `iters` scaffolded Newton iterations of each closed workload
    (optcuts_amd.workloads: cut to a disc, Tutte map, joint-mode a0, first air mesh), the chosen form's solve,
    the air mesh updated between steps as the bench does."""
    from . import workloads as W
    out = []
    ctx = capi.Context(device)
    try:
        for name in names:
            w = W.prepare_closed(ctx, name)
            ctx.newton_init(a0=w["a0"], rel_gl2_tol=1e-12, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG,
                            pcg_rel_tol=pcg_tol, pcg_max_iter=pcg_max_iter)
            for _ in range(iters):
                s, done = ctx.newton_step()
                if s["converged"]:
                    break
                rp, ci, va, b, fm = ctx.get_system()
                out.append((rp, ci, va.copy(), fm))
                if done:
                    break
                ctx.scaffold_update(capi.SCAFFOLD_REMESH_WHEN_NEEDED, 0.0, 2)
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
    ap.add_argument("--closed", default="",
                    help="closed workloads (scaffolded joint-mode systems) to add to the disc systems, e.g. "
                         "icosphere_nu30,tori_h0.06 (measured: lower loss, more PCG iterations on held-out closed "
                         "systems, so the committed weights use discs only)")
    ap.add_argument("--closed-iters", type=int, default=16)
    ap.add_argument("--held-out", default="icosphere_nu38",
                    help="closed workloads whose systems only evaluate the result")
    args = ap.parse_args()
    t0 = time.time()
    meshes = disc_meshes([int(x) for x in args.grids.split(",")], [int(x) for x in args.seeds.split(",")])
    systems = collect_systems(meshes, args.iters)
    n_disc = len(systems)
    if args.closed:
        systems += collect_closed_systems(args.closed.split(","), args.closed_iters)
    print("collected %d systems (%d disc, %d closed) in %.1f s" % (len(systems), n_disc, len(systems) - n_disc,
                                                                   time.time() - t0), flush=True)
    held = collect_closed_systems(args.held_out.split(","), 8) if args.held_out else []
    net = capi.Gnn(0)
    rng = np.random.default_rng(0)
    rhs = [rng.normal(size=2 * (len(s[0]) - 1)) for s in held]
    net.set_params(init_params(0))
    base = [pcg_iterations(net, s, b, tol=1e-6, maxit=5000) for s, b in zip(held, rhs)]
    train(net, systems, args.epochs, lr=args.lr, probes=args.probes, log=lambda m: print(m, flush=True))
    if held:
        got = [pcg_iterations(net, s, b, tol=1e-6, maxit=5000) for s, b in zip(held, rhs)]
        print("held-out closed systems (%s), PCG iterations to 1e-6: untrained (block Jacobi) %s, trained %s; "
              "sum %d -> %d" % (args.held_out, base, got, sum(base), sum(got)), flush=True)
    save(net.get_params(), args.out)
    print("saved %s (%.1f s)" % (args.out, time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
