"""Geometry-step benchmark on one MI355X: Newton iterations/s and time to the reference's final E_SD (1e-6 rel)
at 200k faces, chosen versus plain GPU form.

Headline workload (config C2): three fused tori (genus 3), 200,216 faces, cut to a disc, Tutte map, OptCuts joint
mode's geometry step (energyParams[0] = 1 - 0.999, main.cpp:1569-1570) with the bijectivity scaffold on from the
start (w_scaf = 0.01 energyParams[0], Optimizer.cpp:87). A step is one Newton iteration (Optimizer::solve(1)):
element pass (energy, gradient, PSD-projected Hessian), block-CSR assembly, PCG direction, injectivity bound and
line search, device-resident behind the C ABI, followed by the air-mesh update that the reference does with a new
Triangle mesh (here oc_scaffold_update: rest shape reset, Lawson re-mesh when an air triangle set the step bound
or every 8th step). The timed region starts with the mesh, its Tutte map and the first air mesh in HBM.

    python bench.py [--gpus 1] [--steps 100] [--warmup 1]

Prints one JSON line. Secondary sections: C1 (perturbed icosphere, 100k faces), C4 (icosphere, 1M faces), the
200k height-field disc with all five solver forms, C3 (64 batched charts), the split-query pre-screen.
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
TARGETS = os.path.join(ROOT, "tests", "golden", "workload_targets.json")


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


def element_bytes(nF, nV):
    """Algorithmic bytes of one FP64 element pass (k_element with gradient and projected Hessian) over nF
    triangles on nV vertices: per triangle 3 vertex indices (12 B) and 5 rest features (40 B) read, the gradient
    (3 x 16 B) and the six 2x2 Hessian blocks (6 x 32 B) written; each vertex's UV (16 B) read once."""
    return nF * (12 + 40 + 48 + 192) + 16 * nV


def assembly_bytes(nF, nV, nnzb, nub):
    """Algorithmic bytes of one assembly (k_assemble): per upper block its row, column, two slot positions and
    contribution pointer (20 B); per contribution (6 per triangle: 3 diagonal and 3 off-diagonal element blocks)
    its code (4 B) and its 2x2 block (32 B); every stored block written once (32 B) and per vertex the inverse
    diagonal block (32 B)."""
    return 20 * nub + 36 * 6 * nF + 32 * nnzb + 32 * nV


def cheb_bytes(n, blocks, bs, word=8):
    """Algorithmic bytes of one V-cycle Chebyshev step (k_cheb_step) on a level of n rows and `blocks` stored
    bs x bs blocks: the blocks and their column indices, row pointers, per row the inverse diagonal block, the
    gathered direction (read once), the residual and the iterate read and written, the new direction written."""
    return blocks * (bs * bs * word + 4) + 4 * (n + 1) + n * (bs * bs * word + 6 * bs * word)


def cheb_levels(ctx, reps=50):
    """Per multigrid level with a transfer: rows, blocks, block size, live HIP-event time of one Chebyshev step,
    algorithmic bytes and the rate; and the aggregate over the coarse (3x3-block) levels, whose launches the
    V-cycle makes equally often (two per level per cycle)."""
    levels = ctx.amg_levels()
    out = []
    for l, (dofs, b, bs) in enumerate(levels[:-1]):
        n = dofs // bs
        ms = ctx.time_kernel(10 + l, reps=reps)
        nb = cheb_bytes(n, b, bs)
        out.append({"level": l, "nodes": n, "blocks": b, "block_size": bs, "avg_launch_ms": ms,
                    "algorithmic_bytes": nb, "GB_per_s": nb / (ms * 1e-3) / 1e9,
                    "frac": nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
    coarse = [x for x in out if x["block_size"] == 3]
    agg = None
    if coarse:
        t = sum(x["avg_launch_ms"] for x in coarse)
        nb = sum(x["algorithmic_bytes"] for x in coarse)
        agg = {"avg_launch_ms": t / len(coarse), "algorithmic_bytes_per_launch": nb / len(coarse),
               "achieved": nb / (t * 1e-3) / 1e9}
    return out, agg


# HBM traffic per launch of the roofline kernels, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (KiB per
# dispatch, mean over full dispatches), FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0: the factors that make the
# same counters read the known size of a plain 512 MiB device copy. Summaries with commands:
# profiles/pmc_traffic_200k.md (disc) and profiles/pmc_traffic_tori_200k.md (fused tori).
TRAFFIC = {
    ("disc_200k", "k_lanczos_fine"): (2.0 * 17072.8 + 3188.8) * 1024,
    ("disc_200k", "k_pcg_spmv"): (2.0 * 14334.2 + 3209.3) * 1024,
    # three fused tori, chosen form: mean over multigrid levels 1-3 of the Chebyshev step; the PCG product
    ("tori_200k", "k_cheb_step3"): ((2.0 * 6745.6 + 840.5) + (2.0 * 654.8 + 50.4) + (2.0 * 60.7 + 2.3)) / 3.0 * 1024,
    ("tori_200k", "k_pcg_spmv"): (2.0 * 21640.6 + 3458.0) * 1024,
}

FALLBACK_NAMES = {1: "retry_safe_amg", 2: "jacobi", 3: "truncated", 4: "retry_fp64_elements", 5: "jacobi_direct",
                  6: "truncated_at_cap", 7: "ic0"}


def count_fallbacks(stats):
    out = {v: 0 for v in FALLBACK_NAMES.values()}
    for s in stats:
        if s["pcg_fallback"]:
            out[FALLBACK_NAMES[s["pcg_fallback"]]] += 1
    return out


def load_targets():
    return json.load(open(TARGETS)) if os.path.exists(TARGETS) else {}


# ------------------------------------------------------------------------------------------------------------
# closed surfaces with the scaffold (C1, C2, C4)

def closed_forms(args):
    """name -> (preconditioner, PCG relative tolerance, PCG iteration cap, element precision) on the closed
    surfaces. The fast forms cap PCG per step (inexact Newton; on the Tutte start of a genus-3 map no iterative
    solve captures the exact step and long solves only cost time); the plain form runs Jacobi-PCG to 1e-10."""
    from optcuts_amd import capi
    fp64, mixed = capi.PRECISION_FP64, capi.PRECISION_MIXED
    # (the chosen and mixed forms also loosen the solve to --loose-tol after a step shorter than --loose-below)
    return {"chosen": (capi.PRECOND_AMG, args.chosen_tol, args.chosen_cap, fp64),
            "mixed": (capi.PRECOND_AMG, args.chosen_tol, args.chosen_cap, mixed),
            "learned": (capi.PRECOND_LEARNED, args.closed_learned_tol, 2000, fp64),
            "learned_mixed": (capi.PRECOND_LEARNED, args.closed_learned_tol, 2000, mixed),
            "block_jacobi_at_learned_tol": (capi.PRECOND_BLOCK_JACOBI, args.closed_learned_tol, 2000, fp64),
            "ic0_at_learned_tol": (capi.PRECOND_IC0, args.closed_learned_tol, 2000, fp64),
            "plain": (capi.PRECOND_JACOBI, args.pcg_tol, 200000, fp64)}


def closed_step(ctx, args):
    """One Newton step then the air-mesh update; returns (stats, done, step s, update s, edges flipped). The
    library's set-up phase times of each re-meshing update go to closed_step.remesh_times."""
    from optcuts_amd import capi
    t0 = time.perf_counter()
    s, done = ctx.newton_step()
    t1 = time.perf_counter()
    _, flips = ctx.scaffold_update(capi.SCAFFOLD_REMESH_WHEN_NEEDED, 0.0, args.air_interval)
    t2 = time.perf_counter()
    if flips:
        closed_step.remesh_times.append(ctx.setup_times())
    return s, done, t1 - t0, t2 - t1, flips


closed_step.remesh_times = []


def closed_run(ctx, w, form, args, steps, warmup, e_star=None, budget=None):
    """From the Tutte map: warmup untimed steps, then `steps` timed ones (or, with e_star, until E_SD <= (1+1e-6)
    E* or the budget in seconds). Returns a dict of timings and per-step records."""
    from optcuts_amd import capi
    from optcuts_amd import workloads as W
    pc, tol, mx, prec = closed_forms(args)[form]
    W.restart(ctx, w)
    t0 = time.perf_counter()
    # the forms in --forcing-forms loosen the solve after steps the injectivity bound cut short
    fast_mg = form in args.forcing_forms.split(",")
    ctx.set_forcing(args.loose_tol if fast_mg else 0.0, args.loose_below if fast_mg else 0.0)
    ctx.newton_init(a0=w["a0"], rel_gl2_tol=1e-12, allow_edec_rel_tol=False, precond=pc, pcg_rel_tol=tol,
                    pcg_max_iter=mx, precision=prec)
    t_init = time.perf_counter() - t0
    for _ in range(warmup):
        closed_step(ctx, args)
    rows, stats = [], []
    closed_step.remesh_times = []
    t0 = time.perf_counter()
    hit = None
    k = 0
    while True:
        s, done, ts, tu, nflip = closed_step(ctx, args)
        k += 1
        stats.append(s)
        if getattr(args, "progress", False):
            print("bench: step %d E_SD %.9g pcg %d fallback %d flips %d" % (k, s["E_SD"], s["pcg_iters"],
                                                                          s["pcg_fallback"], nflip),
                  file=sys.stderr, flush=True)
        rows.append((ts, tu, s["E_SD"], s["pcg_iters"], nflip))
        el = time.perf_counter() - t0
        if e_star is not None and hit is None and s["E_SD"] <= (1.0 + 1e-6) * e_star:
            hit = (k, el + t_init)
        if e_star is None and k >= steps:
            break
        if e_star is not None and (hit is not None or k >= steps or (budget and el > budget)):
            break
        if done:
            break
    el = time.perf_counter() - t0
    import numpy as np
    a = np.array(rows)
    out = {"newton_iterations": k, "seconds": el, "newton_iters_per_s": k / el, "ms_per_step": 1e3 * el / k,
           "step_ms_median": float(np.median(a[:, 0]) * 1e3), "air_update_ms_median": float(np.median(a[:, 1]) * 1e3),
           "air_update_share": float(a[:, 1].sum() / el), "init_ms": t_init * 1e3,
           "air_remeshes": int((a[:, 4] > 0).sum()),
           "air_update_ms_median_remesh": float(np.median(a[a[:, 4] > 0, 1]) * 1e3) if (a[:, 4] > 0).any() else None,
           "air_update_ms_median_refresh": float(np.median(a[a[:, 4] == 0, 1]) * 1e3) if (a[:, 4] == 0).any()
           else None,
           "step_ms_max": float(a[:, 0].max() * 1e3), "pcg_iters_max": int(a[:, 3].max()),
           "remesh_phase_ms_median": ({k: float(np.median([t[k] for t in closed_step.remesh_times]))
                                       for k in closed_step.remesh_times[0]} if closed_step.remesh_times else None),
           "remeshes_keeping_amg_structure": sum(1 for t in closed_step.remesh_times
                                                 if t.get("amg_kept_remeshes", 0) > 0),
           "pcg_iters_median": float(np.median(a[:, 3])), "pcg_fallbacks": count_fallbacks(stats),
           "E_SD_first": float(a[0, 2]), "E_SD_last": float(a[-1, 2])}
    if e_star is not None:
        out["reached"] = hit is not None
        if hit is not None:
            out["time_to_target_s"] = hit[1]
            out["iterations_to_target"] = hit[0]
        else:
            out["time_to_target_s"] = None
            out["stopped_after_s"] = el + t_init
        e64 = ctx.energy()
        out["E_SD_reached_fp64"] = e64
        out["rel_err_vs_E_star"] = (e64 - e_star) / e_star
    return out


def closed_bench(ctx, name, args, targets, forms=("chosen", "plain"), timed_steps=None, plain_steps=3,
                 plain_budget=30.0, tte_max=400, form_budget=30.0):
    """The named closed workload: every form's timed steps from the Tutte map, then (with a target) its time to
    E_SD <= (1+1e-6) E*. The plain form runs plain_steps timed steps and its time-to-target stops after
    plain_budget seconds; the other non-chosen forms stop after form_budget seconds."""
    from optcuts_amd import workloads as W
    w = W.prepare_closed(ctx, name)
    e_star = targets.get("closed", {}).get(name, {}).get("E_SD_star")
    out = {"workload": name, "faces": w["meta"]["faces"], "genus": w["meta"]["genus"],
           "cut_vertices": w["meta"]["cut_vertices"], "seam_edges": w["meta"]["seam_edges"],
           "air_triangles": int(w["air"][1].shape[0]), "a0": w["a0"], "w_scaf": w["w_scaf"],
           "generate_and_cut_s": w["generate_s"], "upload_and_tutte_s": w["upload_tutte_s"],
           "tutte_pcg_iters": w["tutte_pcg_iters"], "first_air_mesh_s": w["air_s"], "E_SD_star": e_star,
           "E_SD_star_source": "smallest E_SD of the oracle's exact-Newton run (own LDL^T) from the same Tutte map "
                               "and first air mesh, air rest shape reset after every step, no re-mesh; "
                               "tests/golden/workload_targets.json" if e_star else None, "forms": {}}
    for form in forms:
        print("bench: %s form %s" % (name, form), file=sys.stderr, flush=True)
        pc, tol, mx, prec = closed_forms(args)[form]
        n = min(plain_steps, timed_steps) if form == "plain" else timed_steps
        r = closed_run(ctx, w, form, args, n, args.warmup)
        r.update({"pcg_rel_tol": tol, "pcg_max_iter": mx, "precision": {0: "fp64", 1: "mixed"}[prec],
                  "precond": {0: "jacobi", 1: "block-jacobi", 2: "sa-amg", 3: "learned-factor", 4: "ic0"}[pc]})
        if e_star is not None and not args.no_tte:
            budget = None if form == "chosen" else (plain_budget if form == "plain" else form_budget)
            r["time_to_E_SD_1e-6"] = closed_run(ctx, w, form, args, tte_max, 0, e_star=e_star, budget=budget)
        out["forms"][form] = r
        out["system_sizes"] = ctx.system_sizes()
    tp = out["forms"].get("plain", {}).get("time_to_E_SD_1e-6")
    speed = {}
    for form, r in out["forms"].items():
        tf = r.get("time_to_E_SD_1e-6")
        if not tp or not tf or not tf["reached"]:
            speed[form] = None
        elif tp["reached"]:
            speed[form] = {"speedup": tp["time_to_target_s"] / tf["time_to_target_s"]}
        else:
            speed[form] = {"speedup_at_least": tp["stopped_after_s"] / tf["time_to_target_s"]}
    out["time_to_E_SD_vs_plain"] = speed
    return out, w


# ------------------------------------------------------------------------------------------------------------
# the 200k-face height-field disc (DistMin, no scaffold): all five solver forms

def max_iter_for(precond):
    from optcuts_amd import capi
    return 1000 if precond == capi.PRECOND_AMG else 200000


def run_form(ctx, V, F, UV, pcg_tol, steps, warmup, precond, precision):
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    nV = V.shape[0]
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16 * nV / (nV - 1), allow_edec_rel_tol=False, precond=precond,
                    pcg_rel_tol=pcg_tol, pcg_max_iter=max_iter_for(precond), precision=precision)
    for _ in range(warmup):
        ctx.newton_step()
    t0 = time.perf_counter()
    timed = [ctx.newton_step()[0] for _ in range(steps)]
    return time.perf_counter() - t0, timed


def time_to_target(ctx, V, F, UV, pcg_tol, precond, precision, e_star, max_iter):
    """Wall time of Newton iterations until E_SD <= (1+1e-6) E* (the driver's own energy: FP64 or, in the mixed
    form, FP32 elements with compensated FP64 sums); the reached map is then re-evaluated in FP64."""
    nV = V.shape[0]
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    t0 = time.perf_counter()
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-16 * nV / (nV - 1), allow_edec_rel_tol=False, precond=precond,
                    pcg_rel_tol=pcg_tol, pcg_max_iter=max_iter_for(precond), precision=precision)
    stats, t = [], None
    for k in range(1, max_iter + 1):
        s, done = ctx.newton_step()
        stats.append(s)
        if s["E_SD"] <= (1.0 + 1e-6) * e_star:
            t = time.perf_counter() - t0
            break
        if done:
            break
    e64 = ctx.energy()
    return {"seconds": t, "newton_iterations": k, "E_SD_reached": s["E_SD"], "E_SD_reached_fp64": e64,
            "rel_err_vs_E_star_fp64": abs(e64 - e_star) / e_star, "driver_energy_minus_fp64": s["E_SD"] - e64,
            "pcg_fallbacks": count_fallbacks(stats)}


def disc_bench(ctx, args, targets):
    import numpy as np
    from optcuts_amd import capi, learned
    from optcuts_amd.meshgen import height_field_disc
    V, F, UV = height_field_disc(317, seed=0)
    e_star = targets.get("E_SD_star")
    have_weights = os.path.exists(learned.DEFAULT_WEIGHTS)
    forms = {"plain": (capi.PRECOND_JACOBI, args.pcg_tol, capi.PRECISION_FP64),
             "chosen": (capi.PRECOND_AMG, args.chosen_tol, capi.PRECISION_FP64),
             "mixed": (capi.PRECOND_AMG, args.mixed_tol, capi.PRECISION_MIXED),
             # controls at the loose tolerances of the faster forms: what inexact Newton alone gives
             "plain_at_chosen_tol": (capi.PRECOND_JACOBI, args.chosen_tol, capi.PRECISION_FP64),
             "block_jacobi_at_learned_tol": (capi.PRECOND_BLOCK_JACOBI, args.learned_tol, capi.PRECISION_FP64),
             # the classic incomplete factor on the same pattern, at the learned factor's tolerance
             "ic0_at_learned_tol": (capi.PRECOND_IC0, args.learned_tol, capi.PRECISION_FP64)}
    if have_weights:
        forms["learned"] = (capi.PRECOND_LEARNED, args.learned_tol, capi.PRECISION_FP64)
        forms["learned_mixed"] = (capi.PRECOND_LEARNED, args.learned_tol, capi.PRECISION_MIXED)
    net = None
    if have_weights:
        net = capi.Gnn(0)
        net.set_params(learned.load_weights())
        ctx.set_learned(net)
    if args.disc_forms:
        keep = args.disc_forms.split(",")
        forms = {k: v for k, v in forms.items() if k in keep}
    res = {}
    for name, (precond, tol, precision) in forms.items():
        print("bench: disc_200k form %s" % name, file=sys.stderr, flush=True)
        steps = args.disc_steps if not name.startswith("plain") else min(args.disc_steps, 5)
        el, timed = run_form(ctx, V, F, UV, tol, steps, 1, precond, precision)
        r = {"newton_iters_per_s": steps / el, "ms_per_step": 1e3 * el / steps,
             "pcg_iters_per_newton": float(np.mean([s["pcg_iters"] for s in timed])),
             "pcg_fallbacks": count_fallbacks(timed), "precond": {0: "jacobi", 1: "block-jacobi", 2: "sa-amg",
                                                                   3: "learned-factor", 4: "ic0"}[precond],
             "precision": {0: "fp64", 1: "mixed"}[precision], "pcg_rel_tol": tol}
        r["kernel_ms"] = {"element_pass": ctx.time_kernel(0, reps=10), "assembly": ctx.time_kernel(1, reps=10),
                          "pcg_spmv": ctx.time_kernel(5, reps=50), "pcg_update": ctx.time_kernel(6, reps=50),
                          "linesearch_candidates": ctx.time_kernel(3, reps=10),
                          "step_bound": ctx.time_kernel(4, reps=10)}
        if precond in (capi.PRECOND_AMG, capi.PRECOND_LEARNED, capi.PRECOND_IC0):
            r["kernel_ms"]["precond_apply"] = ctx.time_kernel(7, reps=10)
            r["kernel_ms"]["precond_setup"] = ctx.time_kernel(8, reps=5)
        if precond == capi.PRECOND_AMG:
            r["kernel_ms"]["lanczos_fine_step"] = ctx.time_kernel(9, reps=50)
            if precision == capi.PRECISION_FP64:
                r["cheb_step_per_level"] = cheb_levels(ctx)[0]
        if e_star is not None and not args.no_tte:
            r["time_to_E_SD_1e-6"] = time_to_target(ctx, V, F, UV, tol, precond, precision, e_star, args.tte_max_iter)
        res[name] = r
    if net is not None:
        ctx.set_learned(None)
        net.close()
    nF, nVt, nnzb, nub = ctx.system_sizes()
    speeds = {}
    if "plain" in res and e_star is not None and not args.no_tte:
        tp = res["plain"]["time_to_E_SD_1e-6"]["seconds"]
        for name, r in res.items():
            tf = r["time_to_E_SD_1e-6"]["seconds"]
            speeds[name] = (tp / tf) if (tp and tf) else None
    return {"workload": "height-field disc 317x317, seed 0, 199,712 faces; DistMin (lambda = 0), no scaffold",
            "E_SD_star": e_star, "forms": res, "time_to_E_SD_speedup_vs_plain": speeds,
            "system_sizes": [nF, nVt, nnzb, nub]}, (V, F, UV)


def charts_bench(ctx, max_iter=60):
    """The 64-chart workload (height-field discs of 500-20000 faces, log-uniform sizes, seeds 1000+i): all charts
    batched in one context versus the same charts solved one after another, multigrid form at PCG tolerance 1e-4,
    each chart run until its own stop (gradient tolerance or line-search stop) or max_iter."""
    import numpy as np
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
    # the same batch with FP32 element terms and the FP32-stored multigrid, as many iterations (its charts do
    # not stop on their own at this gradient tolerance, which lies below the FP32 gradient's rounding)
    ctx.set_uv(UV)
    t0 = time.perf_counter()
    ctx.newton_init(precision=capi.PRECISION_MIXED, **kw)
    for it_m in range(1, it + 1):
        s, done = ctx.newton_step()
        if done:
            break
    tm = time.perf_counter() - t0
    st_m = ctx.charts_state()
    t0 = time.perf_counter()
    e_s, its_s = [], 0
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
    e_s = np.array(e_s)
    return {"charts": len(charts), "faces": int(F.shape[0]), "vertices": int(V.shape[0]),
            "batched_seconds": tb, "batched_newton_iterations": it,
            "sequential_seconds": ts, "sequential_newton_iterations_total": its_s,
            "speedup_batched_vs_sequential": ts / tb, "charts_converged": int((st["state"] != 0).sum()),
            "max_rel_diff_final_E_SD": float(abs(e_b - e_s).max() / e_s.min()),
            "batched_mixed_seconds": tm, "batched_mixed_newton_iterations": it_m,
            "batched_mixed_max_rel_diff_final_E_SD": float(abs(st_m["E_SD"] - e_b).max() / e_b.min())}


def prescreen_bench(ctx, V, F, UV, reps=5):
    """Topology-step pre-screen: oc_split_candidates (GPU keys, host selection) for boundary and interior splits,
    median wall time of reps calls."""
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
    """The same pre-screen by the oracle's restatement of computeDivGradPerVert + querySplit's filter (1 thread)."""
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


def cpu_baseline_closed(w):
    """One scaffolded Newton iteration of the headline workload from its Tutte map by the oracle (the project's FP64
    CPU restatement: element terms, assembly, its own sparse LDL^T, line search), 1 thread."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _helpers import ScaffoldOracle
    airUV, Fair, bnd = w["air"]
    t = time.perf_counter()
    so = ScaffoldOracle(w["V"], w["F"], airUV, Fair, bnd, w["area_thres"], w["a0"], w["w_scaf"])
    X = np.concatenate([w["UV0"], airUV[len(bnd):]])
    so.newton_step(X, None)
    el = time.perf_counter() - t
    return {"value": 1.0 / el, "unit": "Newton iters/s", "cores": 1, "kind": "port",
            "sample": "1 scaffolded Newton iteration of the same workload (three fused tori, 200,216 faces, from its "
                      "Tutte map with the same first air mesh): the project's FP64 CPU restatement of the reference "
                      "(own sparse LDL^T, nested-dissection order), 1 thread, %.1f s" % el}


def headline(ctx, args, targets, fast):
    """The headline workload's forms, then the roofline of its dominant kernel: the PCG product (the Jacobi-PCG
    iterations of the steps multigrid cannot precondition and every PCG iteration after them), timed live with
    HIP events on the system after one chosen-form step from the Tutte map."""
    from optcuts_amd import workloads as W
    forms = fast + ("plain",)
    if args.headline_forms:
        forms = tuple(f for f in forms if f in args.headline_forms.split(","))
    print("bench: headline %s" % args.workload, file=sys.stderr, flush=True)
    head, w = closed_bench(ctx, args.workload, args, targets, forms=forms, timed_steps=args.steps,
                           plain_budget=args.plain_budget)
    W.restart(ctx, w)
    pc, tol, mx, prec = closed_forms(args)["chosen"]
    ctx.newton_init(a0=w["a0"], rel_gl2_tol=1e-12, allow_edec_rel_tol=False, precond=pc, pcg_rel_tol=tol,
                    pcg_max_iter=mx, precision=prec)
    closed_step(ctx, args)
    nF, nVt, nnzb, nub = ctx.system_sizes()
    spmv_ms = ctx.time_kernel(5, reps=50)
    nbytes = spmv_bytes(nnzb, nVt)
    achieved = nbytes / (spmv_ms * 1e-3) / 1e9
    spmv = {"bound": "hbm", "kernel": "k_pcg_spmv (PCG product on the 2x2-block Newton matrix, mesh + air)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC.get((args.workload, "k_pcg_spmv")), "algorithmic_bytes_per_launch": nbytes,
            "avg_launch_ms": spmv_ms, "system": {"triangles": nF, "vertices": nVt, "blocks": nnzb}}
    # element pass and assembly of the same Newton system (FP64)
    def rate(kernel, ms, nbytes):
        a = nbytes / (ms * 1e-3) / 1e9
        return {"kernel": kernel, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": ms}
    element = rate("k_element<double,true,true> (energy, gradient, projected Hessian; FP64 VALU-bound)",
                   ctx.time_kernel(0, reps=20), element_bytes(nF, nVt))
    assembly = rate("k_assemble (gather of the element blocks into the block CSR)", ctx.time_kernel(1, reps=20),
                    assembly_bytes(nF, nVt, nnzb, nub))
    # the dominant kernel of the chosen form on this workload: the V-cycle's Chebyshev step on the coarse
    # (3x3-block) levels; per level and over those levels (each launched twice per V-cycle)
    per_level, agg = cheb_levels(ctx)
    roofline = {"bound": "hbm", "kernel": "k_cheb_step<double,3,false> (V-cycle Chebyshev step, coarse levels; "
                                          "aggregate over the levels, which the V-cycle launches equally often)",
                "achieved": agg["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": agg["achieved"] / HBM_PEAK_GBS, "traffic": TRAFFIC.get((args.workload, "k_cheb_step3")),
                "algorithmic_bytes_per_launch": agg["algorithmic_bytes_per_launch"],
                "avg_launch_ms": agg["avg_launch_ms"], "per_level": per_level, "pcg_spmv": spmv,
                "element_pass": element, "assembly": assembly}
    return head, w, roofline


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed Newton steps of the headline chosen form")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tori_200k", choices=["tori_200k", "icosphere_100k", "icosphere_1m"])
    ap.add_argument("--pcg-tol", type=float, default=1e-10, help="plain form PCG relative tolerance")
    ap.add_argument("--chosen-tol", type=float, default=1e-3, help="chosen form (FP64 multigrid) PCG tolerance")
    ap.add_argument("--loose-tol", type=float, default=0.1,
                    help="closed surfaces, chosen and mixed forms: PCG tolerance after a step shorter than "
                         "--loose-below (0: off)")
    ap.add_argument("--loose-below", type=float, default=0.2)
    ap.add_argument("--forcing-forms",
                    default="chosen,mixed,learned,learned_mixed,block_jacobi_at_learned_tol,ic0_at_learned_tol",
                    help="closed-surface forms that use the step-dependent tolerance")
    ap.add_argument("--closed-learned-tol", type=float, default=1e-4,
                    help="closed surfaces: PCG tolerance of the learned forms and their controls (the disc "
                         "section uses --learned-tol)")
    ap.add_argument("--chosen-cap", type=int, default=100,
                    help="PCG iteration cap of the chosen and mixed forms on the closed surfaces")
    ap.add_argument("--mixed-tol", type=float, default=1e-4, help="mixed form PCG tolerance (disc section)")
    ap.add_argument("--learned-tol", type=float, default=5e-3, help="learned forms' PCG tolerance (disc section)")
    ap.add_argument("--air-interval", type=int, default=8,
                    help="re-mesh the air mesh at least every N steps (besides whenever an air triangle set the "
                         "step bound)")
    ap.add_argument("--plain-budget", type=float, default=30.0,
                    help="wall-clock budget (s) of the plain form's time-to-target run on the closed surfaces")
    ap.add_argument("--disc-steps", type=int, default=10)
    ap.add_argument("--tte-max-iter", type=int, default=60)
    ap.add_argument("--no-tte", action="store_true", help="skip the time-to-E_SD runs")
    ap.add_argument("--no-secondary", action="store_true", help="headline workload only")
    ap.add_argument("--no-headline", action="store_true",
                    help="secondary sections only (profiling runs; the JSON line then has no value)")
    ap.add_argument("--headline-forms", default="", help="comma list of the headline's forms (default: all)")
    ap.add_argument("--disc-forms", default="", help="comma list of the disc section's forms (default: all)")
    ap.add_argument("--no-disc", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-charts", action="store_true")
    ap.add_argument("--no-prescreen", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch PCG iterations kernel by kernel")
    ap.add_argument("--progress", action="store_true", help="one stderr line per closed-surface Newton step")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time plain device reads (sum) and copies at several sizes")
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

    from optcuts_amd import capi
    targets = load_targets()
    ctx = capi.Context(0)
    if args.no_graphs:
        ctx.pcg_graphs(False)
    arch = ctx.arch()
    from optcuts_amd import learned
    net = None
    fast = ("chosen", "mixed")
    if os.path.exists(learned.DEFAULT_WEIGHTS):
        net = capi.Gnn(0)
        net.set_params(learned.load_weights())
        ctx.set_learned(net)
        fast = fast + ("learned", "learned_mixed", "block_jacobi_at_learned_tol", "ic0_at_learned_tol")
    from optcuts_amd import workloads as W
    head, w, roofline = None, None, None
    if not args.no_headline:
        head, w, roofline = headline(ctx, args, targets, fast)
    secondary = {}
    if not args.no_secondary:
        if not args.no_c1 and args.workload != "icosphere_100k":
            print("bench: icosphere_100k", file=sys.stderr, flush=True)
            secondary["icosphere_100k"] = closed_bench(ctx, "icosphere_100k", args, targets,
                                                       forms=("chosen", "mixed", "plain"), timed_steps=12,
                                                       plain_steps=5, plain_budget=30.0)[0]
        if not args.no_c4 and args.workload != "icosphere_1m":
            print("bench: icosphere_1m", file=sys.stderr, flush=True)
            secondary["icosphere_1m"] = closed_bench(ctx, "icosphere_1m", args, targets, forms=fast + ("plain",),
                                                     timed_steps=12, plain_steps=2, plain_budget=20.0)[0]
        if not args.no_disc:
            print("bench: disc_200k", file=sys.stderr, flush=True)
            secondary["disc_200k"], (V, F, UV) = disc_bench(ctx, args, targets)
            if not args.no_prescreen:
                secondary["split_prescreen_disc_200k"] = prescreen_bench(ctx, V, F, UV)
                if not args.no_cpu_baseline:
                    secondary["split_prescreen_disc_200k"]["cpu_oracle"] = prescreen_cpu_baseline(V, F, UV)
        if not args.no_charts:
            print("bench: charts", file=sys.stderr, flush=True)
            secondary["batched_charts"] = charts_bench(ctx)
    if net is not None:
        ctx.set_learned(None)
        net.close()
    ctx.close()
    if head is None or "chosen" not in head["forms"]:
        print(json.dumps({"metric": METRIC, "value": None, "unit": "Newton iters/s", "n_gpus": 1,
                          "headline": head, "secondary": secondary}), flush=True)
        return
    chosen = head["forms"]["chosen"]
    out = {
        "metric": METRIC,
        "value": chosen["newton_iters_per_s"],
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%s: %s; cut to a disc, Tutte map, OptCuts joint mode geometry step (lambda_init "
                               "0.999), bijectivity scaffold on; value is the chosen form (FP64, multigrid PCG to "
                               "1e-3) over %d Newton steps from the Tutte map, air-mesh updates included"
                               % (args.workload, W.CLOSED[args.workload], args.steps),
                   "device": arch, "E_SD_star": head["E_SD_star"], "E_SD_star_source": head["E_SD_star_source"]},
        "headline": head,
        "time_to_E_SD_vs_plain": head.get("time_to_E_SD_vs_plain"),
        "secondary": secondary,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_closed(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
