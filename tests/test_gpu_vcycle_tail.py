"""The V-cycle's single-workgroup tail (oc_amg_set_tail): the last smoothed multigrid level and the coarsest
solve run in one launch with the same per-row arithmetic, lane groups and shuffle order as the seven launches it
replaces, so Newton runs with the tail on and off are bit-identical: equal energies, PCG iteration counts and
fallback codes at every step, on the 5k-face disc (FP64 and mixed forms), on the 200k-face disc in the chosen
form (multigrid-PCG to 1e-3) and on 24 batched charts whose coarsest level is block diagonal."""
import pytest

from optcuts_amd import capi
from optcuts_amd.meshgen import height_field_disc

pytestmark = pytest.mark.gpu


def _run(tail, V, F, UV, steps, **kw):
    c = capi.Context(0)
    try:
        c.amg_set_tail(tail)
        c.set_mesh(V, F, fixed=(0,))
        c.set_uv(UV)
        c.newton_init(a0=1.0, rel_gl2_tol=1e-16, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG, **kw)
        levels = c.amg_levels()
        out = []
        for _ in range(steps):
            s, done = c.newton_step()
            out.append((s["E_total"], s["pcg_iters"], s["pcg_fallback"]))
            if done:
                break
        return levels, out
    finally:
        c.close()


@pytest.mark.parametrize("m,steps,precision,tol", [(51, 10, capi.PRECISION_FP64, 1e-6),
                                                    (51, 10, capi.PRECISION_MIXED, 1e-4),
                                                    (317, 5, capi.PRECISION_FP64, 1e-3)])
def test_tail_is_bit_identical(m, steps, precision, tol):
    V, F, UV = height_field_disc(m, seed=0)
    lv_on, on = _run(-1, V, F, UV, steps, pcg_rel_tol=tol, pcg_max_iter=1000, precision=precision)
    lv_off, off = _run(0, V, F, UV, steps, pcg_rel_tol=tol, pcg_max_iter=1000, precision=precision)
    print("levels", lv_on, flush=True)
    for k, (a, b) in enumerate(zip(on, off)):
        print(k, a, b, flush=True)
    assert lv_on == lv_off
    # the tail covers this hierarchy: a 3x3-block last smoothed level of at most 128 nodes above the coarsest
    assert len(lv_on) >= 3 and lv_on[-2][2] == 3 and lv_on[-2][0] // 3 <= 128
    assert len(on) == len(off) and len(on) >= 3
    assert on == off


def test_tail_is_bit_identical_on_batched_charts():
    """24 small charts in one system: coarsening ends at one node per chart, above the dense coarsest size, so
    the coarsest level is block diagonal, the tail's other coarsest-solve path. Per-chart energies are equal
    with the tail on and off at every step."""
    from optcuts_amd.meshgen import union_of_charts
    from _helpers import distort

    charts = []
    for k in range(24):
        V, F, UV = height_field_disc(5, seed=50 + k)
        charts.append((V, F, distort(UV, 70 + k, 0.1)))
    V, F, UV, vs, fs, fixed = union_of_charts(charts)

    def run(tail):
        c = capi.Context(0)
        try:
            c.amg_set_tail(tail)
            c.set_mesh(V, F, fixed=fixed)
            c.set_uv(UV)
            c.set_charts(vs, fs)
            c.newton_init(a0=1.0, rel_gl2_tol=1e-30, allow_edec_rel_tol=False, precond=capi.PRECOND_AMG,
                          pcg_rel_tol=1e-8, pcg_max_iter=10000)
            levels = c.amg_levels()
            out = []
            for _ in range(6):
                s, done = c.newton_step()
                out.append((s["E_total"], s["pcg_iters"], s["pcg_fallback"], tuple(c.charts_state()["E_SD"])))
                if done:
                    break
            return levels, out
        finally:
            c.close()

    lv_on, on = run(-1)
    lv_off, off = run(0)
    print("levels", lv_on, flush=True)
    assert lv_on == lv_off
    rows, blocks, bs = lv_on[-1]
    assert blocks == rows // bs  # block-diagonal coarsest level
    assert len(lv_on) >= 3 and lv_on[-2][2] == 3 and lv_on[-2][0] // 3 <= 128
    assert len(on) >= 3 and on == off
