"""Writes tests/golden/workload_targets.json: the converged energy E_SD* of the
200k-face disc workload (317x317 height-field disc, seed 0) and, under
"closed", of the closed-surface workloads (cut to a disc, Tutte map, joint
objective's geometry step with the scaffold), computed by the CPU oracle (own
FP64 restatement with its own sparse LDL^T), used by bench.py as the targets of
time-to-E_SD(1e-6 rel).

Closed surfaces: exact Newton (the oracle's LDL^T) on a0 E_SD + (w_scaf /
F_air) E_air from the Tutte map with the first air mesh of
optcuts_amd.scaffold; after every step the air mesh keeps its triangles and its
rest shape becomes the current UV (the oracle does not re-mesh), until the
relative change of E_SD over a step is below 1e-9 or max_iter steps. E_SD* is
the smallest E_SD of the run: with the air mesh's rest shape reset after every
step the objective changes slightly from step to step, and on the genus-3
workload E_SD settles into a plateau that moves by about 1e-6 relative between
steps instead of converging to 1e-9.

    python tests/golden/make_workload_targets.py               # disc, ~15 min
    python tests/golden/make_workload_targets.py icosphere_100k # ~10 min
    python tests/golden/make_workload_targets.py tori_200k 45   # ~65 min (85 s per LDL^T step)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import oracle as O  # noqa: E402
from optcuts_amd.meshgen import height_field_disc  # noqa: E402


def main(m=317, seed=0, max_iter=200):
    V, F, UV = height_field_disc(m, seed=seed)
    nV = V.shape[0]
    t = time.time()
    tr, UVf, why = O.newton_run(V, F, UV, a0=1.0, max_iter=max_iter, rel_gl2_tol=1e-16 * nV / (nV - 1),
                                allow_edec_rel_tol=False)
    el = time.time() - t
    out = {
        "workload": "height_field_disc m=%d seed=%d (%d faces, %d vertices)" % (m, seed, F.shape[0], nV),
        "E_SD_star": float(tr[-1, 1]),
        "newton_iterations": int(len(tr) - 1),
        "exit_reason": int(why),
        "E_SD_trace": [float(x) for x in tr[:, 1]],
        "oracle_seconds": el,
    }
    path = os.path.join(HERE, "workload_targets.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, out["E_SD_star"], out["newton_iterations"], el)


def closed(name, max_iter=60, lam=0.999):
    sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..")))
    from _helpers import ScaffoldOracle
    from optcuts_amd import scaffold as SC
    from optcuts_amd import surfaces as S
    Vc, Fc, cohE, meta = S.workload(name)
    UV = O.tutte_uv(Vc, Fc)[0]
    ael = O.RestFeatures(Vc, Fc).avgEdgeLen
    thr = SC.area_threshold(ael)
    airUV, Fair, bnd = SC.initial_air_mesh(UV, Fc, ael)
    a0 = 1.0 - lam
    X = np.concatenate([UV, airUV[len(bnd):]])
    trace = []
    t = time.time()
    for k in range(max_iter):
        so = ScaffoldOracle(Vc, Fc, airUV, Fair, bnd, thr, a0, 0.01 * a0)
        X, E, step, halv, gsq = so.newton_step(X, None)
        airUV = np.concatenate([X[bnd], X[Vc.shape[0]:]])
        trace.append(float(so.energy(X)[1]))
        print(name, k, trace[-1], step, flush=True)
        if k > 0 and abs(trace[-1] - trace[-2]) <= 1e-9 * trace[-1]:
            break
    path = os.path.join(HERE, "workload_targets.json")
    d = json.load(open(path))
    d.setdefault("closed", {})[name] = {"workload": meta, "E_SD_star": min(trace), "E_SD_last": trace[-1],
                                        "newton_iterations": len(trace), "E_SD_trace": trace,
                                        "oracle_seconds": time.time() - t, "lambda_init": lam}
    with open(path, "w") as f:
        json.dump(d, f, indent=1)
    print("wrote", path, name, min(trace), len(trace))


if __name__ == "__main__":
    if len(sys.argv) > 1:
        closed(sys.argv[1], *[int(a) for a in sys.argv[2:3]])
    else:
        main()
