"""Set-up of the geometry-step workloads on a device context.

Closed surfaces (perturbed icospheres, fused tori) follow the reference's start for a mesh without input UV
(main.cpp:1380-1532): cut to a disc, Tutte map with the boundary on the unit circle, then the joint OptCuts
objective's geometry step: energyParams[0] = 1 - lambda_init (main.cpp:1569-1570; lambda_init = 0.999 in joint
mode), vertex 0 fixed (TriMesh.cpp:325-328), the bijectivity scaffold on from the start (initCut 1, main.cpp:1572)
with w_scaf = 0.01 energyParams[0] (Optimizer.cpp:87) and areaThres_AM from avgEdgeLen / 2 (Scaffold.cpp:34,176).
"""
import time

import numpy as np

from optcuts_amd import capi
from optcuts_amd import scaffold as SC
from optcuts_amd import surfaces as S
from optcuts_amd.meshgen import height_field_disc

CLOSED = {
    # name: (description, generator)
    "icosphere_100k": "perturbed icosphere, geodesic nu=71 (100,820 faces), 6 lobes width U(0.3,0.8), seed 1",
    "tori_200k": "three fused tori (R=1, r=0.35, centres x=-1.8,0,1.8), surface nets h=0.0219 + relaxation "
                 "(200,216 faces, genus 3), jitter 0.01 edge, seed 2",
    "icosphere_1m": "perturbed icosphere, geodesic nu=224 (1,003,520 faces), 6 lobes width U(0.2,0.6), seed 4",
}


def prepare_closed(ctx, name, lam=0.999, tutte_precond=capi.PRECOND_AMG, scaffold=True):
    """Cut the named closed surface (optcuts_amd.surfaces.workload), upload it, compute the Tutte map on the GPU
    and attach the initial air mesh. Returns a dict with the arrays and parameters (a0, w_scaf, area_thres) and
    the host/GPU set-up times."""
    t0 = time.perf_counter()
    Vc, Fc, cohE, meta = S.workload(name)
    t1 = time.perf_counter()
    ctx.set_mesh(Vc, Fc, fixed=(0,))
    UV, nl, nb, it = ctx.tutte(capi.TUTTE_COMPONENTS, precond=tutte_precond, rel_tol=1e-12, max_iter=200000)
    t2 = time.perf_counter()
    _, ael = ctx.mesh_stats()
    a0 = 1.0 - lam
    out = {"V": Vc, "F": Fc, "cohE": cohE, "UV0": UV, "meta": meta, "a0": a0, "w_scaf": 0.01 * a0,
           "avg_edge_len": ael, "area_thres": SC.area_threshold(ael), "tutte_pcg_iters": it,
           "generate_s": t1 - t0, "upload_tutte_s": t2 - t1}
    if scaffold:
        air = SC.initial_air_mesh(UV, Fc, ael)
        ctx.set_air_mesh(*air, out["area_thres"], out["w_scaf"])
        out["air"] = air
        out["air_s"] = time.perf_counter() - t2
    return out


def restart(ctx, w, scaffold=True):
    """Back to the initial map (and initial air mesh) of a prepared workload."""
    ctx.set_mesh(w["V"], w["F"], fixed=(0,))
    ctx.set_uv(w["UV0"])
    if scaffold and "air" in w:
        ctx.set_air_mesh(*w["air"], w["area_thres"], w["w_scaf"])


def prepare_disc(ctx, m=317, seed=0, scaffold=False, a0=1.0):
    """Height-field disc (DistMin, lambda = 0) with the xy projection as its map."""
    V, F, UV = height_field_disc(m, seed=seed)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    _, ael = ctx.mesh_stats()
    out = {"V": V, "F": F, "UV0": UV, "a0": a0, "w_scaf": 0.01 * a0, "avg_edge_len": ael,
           "area_thres": SC.area_threshold(ael)}
    if scaffold:
        air = SC.initial_air_mesh(UV, F, ael)
        ctx.set_air_mesh(*air, out["area_thres"], out["w_scaf"])
        out["air"] = air
    return out
