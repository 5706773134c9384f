"""Geometry-step benchmark: Newton iterations/s and time-to-E_SD(1e-6 rel) on
the 200k-face synthetic workload (317x317 height-field disc, seed 0,
199,712 faces), one MI355X.

A step is one Newton iteration of the symmetric Dirichlet geometry step
(Optimizer::solve(1)): element pass (energy, gradient, PSD-projected
Hessian), block-CSR assembly, PCG direction, injectivity step bound and line
search, all device-resident behind the C ABI.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]

Prints one JSON line. The timed region starts with the mesh, UV and rest
features already in HBM.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newton iters/sec + time-to-E_SD(1e-6 rel) at 200k faces; chosen vs plain GPU"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(nnzb, nV):
    """Algorithmic bytes of one PCG SpMV launch on the 2x2-block CSR:
    values 32 B + column index 4 B per stored block, row pointers 4 B per row,
    z and p_old read once (16 B each per vertex), p_new and q written once."""
    return 36 * nnzb + 4 * (nV + 1) + 64 * nV


def lanczos0_bytes(nnzb, nV):
    """Algorithmic bytes of one fine-level Lanczos step of the multigrid set-up
    (k_lanczos on the 2x2-block CSR): values 32 B + column index 4 B per block
    (the diagonal block, re-read for D and its inverse, is one of them), row
    pointers 4 B per row, per vertex the diagonal position (4 B), W_{j-1},
    W_{j-2}, v_{j-1}, v_{j-2} read once (64 B) and v_j, W_j written (32 B)."""
    return 36 * nnzb + 4 * (nV + 1) + 100 * nV


# HBM traffic per launch of the roofline kernels at the 200k-face workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch, mean over
# full dispatches: those queued past PCG convergence return at once and are
# left out), FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0: the factors that
# make the same counters read the known size of a plain 512 MiB device copy.
# Summaries with commands: profiles/pmc_traffic_200k.md.
TRAFFIC_200K = {
    "k_lanczos_fine": (2.0 * 17072.8 + 3188.8) * 1024,
    "k_pcg_spmv": (2.0 * 14334.2 + 3209.3) * 1024,
}


def max_iter_for(precond):
    from optcuts_amd import capi
    return 1000 if precond == capi.PRECOND_AMG else 200000


def run_form(ctx, V, F, UV, pcg_tol, steps, warmup, precond, precision):
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    nV = V.shape[0]
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16 * nV / (nV - 1), allow_edec_rel_tol=False, precond=precond,
                    pcg_rel_tol=pcg_tol, pcg_max_iter=max_iter_for(precond), precision=precision)
    stats = []
    for _ in range(warmup):
        s, done = ctx.newton_step()
        stats.append(s)
    t0 = time.perf_counter()
    timed = []
    for _ in range(steps):
        s, done = ctx.newton_step()
        timed.append(s)
    t1 = time.perf_counter()
    return t1 - t0, stats, timed


def time_to_target(ctx, V, F, UV, pcg_tol, precond, precision, e_star, max_iter):
    """Wall time of Newton iterations until |E_SD - E*| <= 1e-6 E* (the driver's
    own energy: FP64 or, in the mixed form, FP32 elements with compensated
    FP64 sums). The reached map is then re-evaluated in FP64 and that value
    is checked too, so the mixed form's energy error is stated per result."""
    nV = V.shape[0]
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    t0 = time.perf_counter()
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16 * nV / (nV - 1), allow_edec_rel_tol=False, precond=precond,
                    pcg_rel_tol=pcg_tol, pcg_max_iter=max_iter_for(precond), precision=precision)
    fb = {"retry_fp64_elements": 0, "retry_safe_amg": 0, "jacobi": 0, "truncated": 0}
    t = None
    fixed = e_star is None  # no target: run max_iter iterations, the reached E_SD becomes the target
    for k in range(1, max_iter + 1):
        s, done = ctx.newton_step()
        if s["pcg_fallback"] == 4:
            fb["retry_fp64_elements"] += 1
        elif s["pcg_fallback"] == 1:
            fb["retry_safe_amg"] += 1
        elif s["pcg_fallback"] == 2:
            fb["jacobi"] += 1
        elif s["pcg_fallback"] == 3:
            fb["truncated"] += 1
        if not fixed and abs(s["E_SD"] - e_star) <= 1e-6 * e_star:
            t = time.perf_counter() - t0
            break
        if done:
            break
    if fixed:
        t = time.perf_counter() - t0
        e_star = s["E_SD"]
    e64 = ctx.energy()
    return {"seconds": t, "newton_iterations": k, "E_SD_reached": s["E_SD"], "E_SD_reached_fp64": e64,
            "rel_err_vs_E_star_fp64": abs(e64 - e_star) / e_star,
            "driver_energy_minus_fp64": s["E_SD"] - e64, "pcg_fallbacks": fb}


def charts_bench(ctx, max_iter=60):
    """The 64-chart workload (height-field discs of 500-20000 faces, log-uniform
    sizes, seeds 1000+i): all charts batched in one context versus the same
    charts solved one after another, multigrid form at PCG tolerance 1e-4, each chart run until
    its own stop (gradient tolerance or line-search stop) or max_iter."""
    from optcuts_amd import capi
    from optcuts_amd.meshgen import batched_disc_charts, union_of_charts
    charts = batched_disc_charts(64, 500, 20000, base_seed=1000)
    V, F, UV, vs, fs, fixed = union_of_charts(charts)
    kw = dict(a0=1.0, rel_gl2_tol=1e-16, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG, pcg_rel_tol=1e-4,
              pcg_max_iter=1000)
    ctx.set_mesh(V, F, fixed=fixed)
    ctx.set_uv(UV)
    ctx.set_charts(vs, fs)
    t0 = time.perf_counter()
    ctx.newton_init(**kw)
    for it in range(1, max_iter + 1):
        s, done = ctx.newton_step()
        if done:
            break
    tb = time.perf_counter() - t0
    st = ctx.charts_state()
    e_b = st["E_SD"].copy()
    t0 = time.perf_counter()
    e_s = []
    its_s = 0
    for Vk, Fk, UVk in charts:
        ctx.set_mesh(Vk, Fk, fixed=(0,))
        ctx.set_uv(UVk)
        ctx.newton_init(**kw)
        for k in range(1, max_iter + 1):
            s, done = ctx.newton_step()
            if done:
                break
        its_s += k
        e_s.append(s["E_SD"])
    ts = time.perf_counter() - t0
    e_s = __import__("numpy").array(e_s)
    return {"charts": len(charts), "faces": int(F.shape[0]), "vertices": int(V.shape[0]),
            "batched_seconds": tb, "batched_newton_iterations": it,
            "sequential_seconds": ts, "sequential_newton_iterations_total": its_s,
            "speedup_batched_vs_sequential": ts / tb,
            "charts_converged": int((st["state"] != 0).sum()),
            "max_rel_diff_final_E_SD": float(abs(e_b - e_s).max() / e_s.min())}


def scaffold_bench(ctx, V, F, UV, forms, iters=10, w_scaf=0.01, e_star=None, max_iter=40):
    """The geometry loop with the bijectivity scaffold, as the reference runs it by default: before every
    Newton step a new air mesh is made at the current UV (Optimizer.cpp:251-257; the reference re-meshes the
    air region with Triangle, here a two-row band outside the boundary with fresh random diagonals stands in
    for it), set (oc_set_air_mesh), and the system re-initialised (new pattern, slot maps and multigrid
    structure: oc_newton_init), then one step. Per-iteration wall times of the library calls; the air-mesh
    generation is timed apart and left out of iters_per_s and of the time to target. With e_star, each form
    runs until its mesh E_SD is within 1e-6 of it (at most max_iter iterations), else iters iterations."""
    import numpy as np
    from optcuts_amd.meshgen import band_air_mesh, boundary_loop_ccw
    loop = boundary_loop_ccw(F)
    e = np.linalg.norm(UV[F[:, [1, 2, 0]]] - UV[F], axis=2).mean()
    area_thres = np.sqrt(3.0) / 4.0 * (0.5 * e) ** 2
    out = {}
    for name, (precond, tol) in forms.items():
        ctx.set_mesh(V, F, fixed=(0,))
        ctx.set_uv(UV)
        rows = []
        hit = None
        for k in range(max_iter if e_star is not None else iters):
            t0 = time.perf_counter()
            airUV, Fair, bnd = band_air_mesh(ctx.get_uv(), loop, rings=2, seed=k)
            t1 = time.perf_counter()
            ctx.set_air_mesh(airUV, Fair, bnd, area_thres, w_scaf)
            ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16, allow_edec_rel_tol=False, precond=precond, pcg_rel_tol=tol,
                            pcg_max_iter=max_iter_for(precond))
            t2 = time.perf_counter()
            s, done = ctx.newton_step()
            t3 = time.perf_counter()
            rows.append((t1 - t0, t2 - t1, t3 - t2, s["pcg_iters"], s["pcg_fallback"], s["E_SD"]))
            if e_star is not None and hit is None and abs(s["E_SD"] - e_star) <= 1e-6 * e_star:
                hit = len(rows)
            if done or (hit is not None and len(rows) >= iters):
                break
        a = np.array([r[:3] for r in rows]) * 1e3
        out[name] = {"newton_iterations": len(rows), "pcg_rel_tol": tol,
                     "iters_per_s": len(rows) / (a[:, 1].sum() + a[:, 2].sum()) * 1e3,
                     "reinit_ms_median": float(np.median(a[1:, 1])) if len(rows) > 1 else float(a[0, 1]),
                     "step_ms_median": float(np.median(a[1:, 2])) if len(rows) > 1 else float(a[0, 2]),
                     "first_iteration_ms": float(a[0, 1] + a[0, 2]),
                     "air_mesh_generation_ms_median": float(np.median(a[:, 0])),
                     "pcg_iters": [int(r[3]) for r in rows], "pcg_fallback_codes": [int(r[4]) for r in rows],
                     "E_SD_last": rows[-1][5],
                     "air_triangles": int(Fair.shape[0]), "air_vertices": int(airUV.shape[0])}
        if e_star is not None:
            out[name]["time_to_E_SD_1e-6"] = (
                {"seconds": float((a[:hit, 1].sum() + a[:hit, 2].sum()) * 1e-3), "newton_iterations": hit}
                if hit else None)
    if "chosen" in out and "plain" in out:
        out["speedup_iters_per_s_chosen_vs_plain"] = out["chosen"]["iters_per_s"] / out["plain"]["iters_per_s"]
        tc, tp = out["chosen"].get("time_to_E_SD_1e-6"), out["plain"].get("time_to_E_SD_1e-6")
        if tc and tp:
            out["time_to_E_SD_speedup_chosen_vs_plain"] = tp["seconds"] / tc["seconds"]
    return out


def prescreen_bench(ctx, V, F, UV, reps=5):
    """Topology-step pre-screen at the workload's size: oc_split_candidates (GPU keys, host selection) for
    boundary and interior splits, median wall time of reps calls."""
    import numpy as np
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    out = {}
    for name, interior in (("boundary", False), ("interior", True)):
        ctx.split_candidates(interior)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            cand, ns = ctx.split_candidates(interior)
            ts.append(time.perf_counter() - t0)
        out[name] = {"ms_median": float(np.median(ts) * 1e3), "candidates": int(len(cand)), "map_size": int(ns)}
    return out


def prescreen_cpu_baseline(V, F, UV):
    """The same pre-screen by the oracle's restatement of computeDivGradPerVert + querySplit's filter
    (1 thread)."""
    from oracle import oracle as O
    feat = O.RestFeatures(V, F)
    out = {}
    for name, interior in (("boundary", False), ("interior", True)):
        t0 = time.perf_counter()
        dg = O.div_grad_per_vert(F, UV, feat)
        O.split_candidates(F, V.shape[0], dg, None, interior, 0.6)
        out[name + "_ms"] = (time.perf_counter() - t0) * 1e3
    out["cores"] = 1
    return out


def cpu_baseline(V, F, UV, iters=3):
    from oracle import oracle as O
    t = time.perf_counter()
    tr, _, _ = O.newton_run(V, F, UV, a0=1.0, max_iter=iters, rel_gl2_tol=1e-30, allow_edec_rel_tol=False)
    el = time.perf_counter() - t
    return {"value": (len(tr) - 1) / el, "unit": "Newton iters/s", "cores": 1, "kind": "port",
            "sample": "%d Newton iterations of the same 200k-face workload from the same initial UV: the "
                      "project's FP64 CPU restatement of the reference (own sparse LDL^T, nested-dissection "
                      "order), 1 thread, %.1f s" % (len(tr) - 1, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=317, help="vertex grid side (317 -> 199,712 faces)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--pcg-tol", type=float, default=1e-10, help="plain form PCG relative tolerance")
    ap.add_argument("--chosen-tol", type=float, default=1e-3,
                    help="chosen form (FP64 multigrid) PCG relative tolerance")
    ap.add_argument("--mixed-tol", type=float, default=1e-4,
                    help="mixed form PCG relative tolerance (at 1e-3 its FP32-stored hierarchy falls back to "
                         "Jacobi on this workload)")
    ap.add_argument("--learned-tol", type=float, default=5e-3,
                    help="PCG relative tolerance of both learned forms (a factor that needs hundreds of PCG "
                         "iterations for 1e-4 reaches the target sooner with looser inner solves)")
    ap.add_argument("--form", choices=["all", "both", "plain", "chosen", "mixed", "learned", "learned_mixed"],
                    default="all", help="all = every form (learned ones when trained weights exist); "
                                        "both = plain and chosen")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-charts", action="store_true", help="skip the 64-chart batched run")
    ap.add_argument("--no-tte", action="store_true", help="skip the time-to-E_SD run")
    ap.add_argument("--no-scaffold", action="store_true", help="skip the scaffolded loop")
    ap.add_argument("--no-prescreen", action="store_true", help="skip the split-query pre-screen timing")
    ap.add_argument("--no-graphs", action="store_true", help="launch PCG iterations kernel by kernel")
    ap.add_argument("--scaffold-iters", type=int, default=10)
    ap.add_argument("--tte-max-iter", type=int, default=60)
    ap.add_argument("--target-iters", type=int, default=0, metavar="K",
                    help="workloads without a stored E*: the plain form's E_SD after K Newton iterations from "
                         "the initial map is the time-to-target goal of the other forms")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time plain device reads (sum) and copies at several sizes: what the GPU streams at "
                         "the row kernels' size, for comparison with their roofline fraction")
    args = ap.parse_args()
    if args.stream_probe:
        import torch
        out = {}
        for mb in (16, 36, 72, 144, 512):
            x = torch.ones(mb * 1000000 // 8, dtype=torch.float64, device="cuda")
            y = torch.empty_like(x)
            r = torch.empty((), dtype=torch.float64, device="cuda")
            for name, fn, nbytes in (("read_sum", lambda: torch.sum(x, dim=0, out=r), x.numel() * 8),
                                     ("copy", lambda: y.copy_(x), 2 * x.numel() * 8)):
                for _ in range(5):
                    fn()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(50):
                    fn()
                e1.record()
                torch.cuda.synchronize()
                us = e0.elapsed_time(e1) * 1e3 / 50
                out["%s_%dMB" % (name, mb)] = {"us": us, "GB_per_s": nbytes / (us * 1e-6) / 1e9}
        print(json.dumps({"stream_probe": out}), flush=True)
        return
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_per_launch": a.numel() * 8}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: the geometry step of one chart is a single-GPU workload in this "
              "build (multi-GPU sharding is a design note only); got --gpus %d, WORLD_SIZE %d" % (args.gpus, world),
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from optcuts_amd import capi
    from optcuts_amd.meshgen import height_field_disc

    V, F, UV = height_field_disc(args.grid, seed=args.seed)
    forms = {"plain": (capi.PRECOND_JACOBI, args.pcg_tol, capi.PRECISION_FP64),
             "chosen": (capi.PRECOND_AMG, args.chosen_tol, capi.PRECISION_FP64),
             "mixed": (capi.PRECOND_AMG, args.mixed_tol, capi.PRECISION_MIXED),
             "learned": (capi.PRECOND_LEARNED, args.learned_tol, capi.PRECISION_FP64),
             "learned_mixed": (capi.PRECOND_LEARNED, args.learned_tol, capi.PRECISION_MIXED)}
    from optcuts_amd import learned
    have_weights = os.path.exists(learned.DEFAULT_WEIGHTS)
    if args.form == "all":
        run = ["plain", "chosen", "mixed"] + (["learned", "learned_mixed"] if have_weights else [])
    elif args.form == "both":
        run = ["plain", "chosen"]
    else:
        run = [args.form]
    ctx = capi.Context(0)
    if args.no_graphs:
        ctx.pcg_graphs(False)
    net = None
    if any(f.startswith("learned") for f in run):
        net = capi.Gnn(0)
        net.set_params(learned.load_weights())
        ctx.set_learned(net)
    arch = ctx.arch()
    targets_path = os.path.join(ROOT, "tests", "golden", "workload_targets.json")
    e_star = None
    if os.path.exists(targets_path) and args.grid == 317 and args.seed == 0:
        e_star = json.load(open(targets_path))["E_SD_star"]
    e_star_source = "oracle run to convergence (tests/golden/workload_targets.json)" if e_star is not None else None
    if e_star is None and args.target_iters > 0 and "plain" in run:
        run = ["plain"] + [f for f in run if f != "plain"]
    res = {}
    for name in run:
        precond, tol, precision = forms[name]
        el, warm, timed = run_form(ctx, V, F, UV, tol, args.steps, args.warmup, precond, precision)
        r = {"newton_iters_per_s": args.steps / el, "ms_per_step": 1e3 * el / args.steps,
             "pcg_iters_per_newton": float(np.mean([s["pcg_iters"] for s in timed])),
             "pcg_fallback_codes": [s["pcg_fallback"] for s in timed],
             "precond_setups": [s["precond_setups"] for s in timed],
             "precond": {capi.PRECOND_JACOBI: "jacobi", capi.PRECOND_AMG: "sa-amg",
                         capi.PRECOND_LEARNED: "learned-factor"}[precond],
             "precision": {capi.PRECISION_FP64: "fp64", capi.PRECISION_MIXED: "mixed"}[precision],
             "pcg_rel_tol": tol}
        nF, nVt, nnzb, nub = ctx.system_sizes()
        r["kernel_ms"] = {"element_pass": ctx.time_kernel(0, reps=10), "assembly": ctx.time_kernel(1, reps=10),
                          "pcg_spmv": ctx.time_kernel(5, reps=50), "pcg_update": ctx.time_kernel(6, reps=50),
                          "linesearch_candidates": ctx.time_kernel(3, reps=10),
                          "step_bound": ctx.time_kernel(4, reps=10)}
        if precond in (capi.PRECOND_AMG, capi.PRECOND_LEARNED):
            r["kernel_ms"]["precond_apply"] = ctx.time_kernel(7, reps=10)
            r["kernel_ms"]["precond_setup"] = ctx.time_kernel(8, reps=5)
        if precond == capi.PRECOND_AMG:
            r["kernel_ms"]["lanczos_fine_step"] = ctx.time_kernel(9, reps=50)
        if e_star is None and name == "plain" and args.target_iters > 0 and not args.no_tte:
            r["time_to_E_SD_1e-6"] = time_to_target(ctx, V, F, UV, tol, precond, precision, None,
                                                    args.target_iters)
            e_star = r["time_to_E_SD_1e-6"]["E_SD_reached"]
            e_star_source = "plain form after %d Newton iterations" % args.target_iters
        elif e_star is not None and not args.no_tte:
            r["time_to_E_SD_1e-6"] = time_to_target(ctx, V, F, UV, tol, precond, precision, e_star,
                                                    args.tte_max_iter)
        res[name] = r
    charts = None if args.no_charts else charts_bench(ctx)
    scaf = None
    if not args.no_scaffold:
        sf = {}
        if "chosen" in res:
            sf["chosen"] = (capi.PRECOND_AMG, args.chosen_tol)
        if "plain" in res:
            sf["plain"] = (capi.PRECOND_JACOBI, args.pcg_tol)
        scaf = scaffold_bench(ctx, V, F, UV, sf, iters=args.scaffold_iters, e_star=e_star) if sf else None
    pre = None if args.no_prescreen else prescreen_bench(ctx, V, F, UV)
    ctx.close()
    if net is not None:
        net.close()

    head = res["chosen"] if "chosen" in res else (res["plain"] if "plain" in res else next(iter(res.values())))
    plain = res.get("plain")
    is_200k = args.grid == 317 and args.seed == 0

    def roofline(kernel, nbytes, ms, traffic_key):
        achieved = nbytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_200K[traffic_key] if is_200k else None,
                "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": ms}

    if "chosen" in res:
        # the chosen form's dominant kernel (kernel trace: k_lanczos, fine level)
        rl = roofline("k_lanczos<double,2> (fine-level Lanczos step of the multigrid set-up)",
                      lanczos0_bytes(nnzb, nVt), res["chosen"]["kernel_ms"]["lanczos_fine_step"], "k_lanczos_fine")
    else:
        rl = roofline("k_pcg_spmv", spmv_bytes(nnzb, nVt), head["kernel_ms"]["pcg_spmv"], "k_pcg_spmv")
    rl_plain = roofline("k_pcg_spmv", spmv_bytes(nnzb, nVt), plain["kernel_ms"]["pcg_spmv"],
                        "k_pcg_spmv") if plain else None
    speeds = {}
    if "plain" in res and e_star is not None and not args.no_tte:
        tp = res["plain"]["time_to_E_SD_1e-6"]["seconds"]
        for name, r in res.items():
            tf = r["time_to_E_SD_1e-6"]["seconds"]
            speeds[name] = (tp / tf) if (tp and tf) else None
    speed = speeds.get("chosen")
    out = {
        "metric": METRIC,
        "value": head["newton_iters_per_s"],
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "height-field disc %dx%d grid, seed %d, %d faces, %d vertices; DistMin (lambda=0), "
                               "no scaffold; value is the %s form" % (args.grid, args.grid, args.seed, F.shape[0],
                                                                      V.shape[0], "chosen" if "chosen" in res
                                                                      else "plain"),
                   "device": arch, "E_SD_star": e_star, "E_SD_star_source": e_star_source},
        "forms": res,
        "time_to_E_SD_speedup_chosen_vs_plain": speed,
        "time_to_E_SD_speedup_vs_plain": speeds,
        "batched_charts": charts,
        "scaffold": scaf,
        "split_prescreen": pre,
        "roofline": rl,
        "roofline_plain_form": rl_plain,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(V, F, UV)
        if pre is not None:
            pre["cpu_oracle"] = prescreen_cpu_baseline(V, F, UV)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
