"""Writes tests/golden/workload_targets.json: the converged energy E_SD* of the
200k-face benchmark workload (317x317 height-field disc, seed 0), computed by
the CPU oracle (own FP64 restatement with its own sparse LDL^T), used by
bench.py as the target of time-to-E_SD(1e-6 rel).

Runs about 15 minutes on one core:  python tests/golden/make_workload_targets.py
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


if __name__ == "__main__":
    main()
