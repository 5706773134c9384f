"""Per-iteration trace files of the device Newton loop (energyValPerIter.txt,
gradientPerIter.txt; Optimizer.cpp:61-62, 196-214, 645-648, 706-739): one energy
line at oc_newton_init and after every step that did not stop, one gradient line
at the start of every step, in the default std::ostream number format ("%g",
pinned against the stream in tests/test_io.py)."""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd import capi
from optcuts_amd.meshgen import height_field_disc

from _helpers import distort, grid_air_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def _read(path):
    with open(path) as f:
        return f.read().split("\n")[:-1]


def test_traces_without_scaffold(ctx, tmp_path):
    V, F, UV = height_field_disc(21, seed=5)
    UVd = distort(UV, 5, 0.05)
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UVd)
    a0, seam = 1.5, 0.25
    esd0 = ctx.energy()
    ctx.trace_open(str(tmp_path))
    ctx.trace_set_seam(seam)
    E0 = ctx.newton_init(a0=a0, rel_gl2_tol=1e-12, allow_edec_rel_tol=True, pcg_rel_tol=1e-12, pcg_max_iter=20000)
    elines = ["%g %g %g %g" % (E0, esd0, seam, a0)]
    glines = []
    for _ in range(60):
        s, done = ctx.newton_step()
        glines.append("%g %g" % (s["grad_sq"], s["grad_sq"] / (a0 * a0)))
        if not s["stopped"] and not s["converged"]:
            elines.append("%g %g %g %g" % (s["E_total"], s["E_SD"], seam, a0))
        if done:
            break
    assert done
    ctx.trace_close()
    assert _read(tmp_path / "energyValPerIter.txt") == elines
    assert _read(tmp_path / "gradientPerIter.txt") == glines
    assert len(glines) >= 3


def test_traces_with_scaffold_report_the_mesh_gradient(ctx, tmp_path):
    """With an air mesh the second gradient column is the mesh SD term's own |g|^2
    (Optimizer.cpp:731-734), not the total's."""
    m = 11
    V, F, UV = height_field_disc(m, seed=3)
    airUV, Fair, bnd = grid_air_mesh(m, rings=2)
    area_thres = np.sqrt(3.0) / 4.0 * (0.5 * O.RestFeatures(V, F).avgEdgeLen) ** 2
    ctx.set_mesh(V, F, fixed=(0,))
    ctx.set_uv(UV)
    ctx.set_air_mesh(airUV, Fair, bnd, area_thres, 1.0)  # a heavy air term, so it shows in |g|^2
    ctx.trace_open(str(tmp_path))
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-30, allow_edec_rel_tol=False, pcg_rel_tol=1e-12, pcg_max_iter=20000)
    gsd = []
    gtot = []
    for _ in range(3):
        g = ctx.gradient()
        gsd.append(float(g @ g))
        s, _ = ctx.newton_step()
        gtot.append(s["grad_sq"])
    ctx.trace_close()
    rows = [[float(x) for x in l.split()] for l in _read(tmp_path / "gradientPerIter.txt")]
    assert len(rows) == 3
    for r, a, b in zip(rows, gtot, gsd):
        assert r[0] == pytest.approx(a, rel=1e-5)
        assert r[1] == pytest.approx(b, rel=1e-5)
    # the air mesh starts at its rest shape (no air gradient); once the first step has
    # deformed it, the total and the mesh term differ
    assert all(abs(a - b) > 1e-6 * a for a, b in zip(gtot[1:], gsd[1:])), (gtot, gsd)
    assert len(_read(tmp_path / "energyValPerIter.txt")) == 4


def test_obj_to_tutte_to_newton_to_obj(ctx, tmp_path):
    """The geometry step end to end from files: a disc read from OBJ, Tutte's initial map,
    Newton with the multigrid preconditioner to the gradient tolerance, the result written
    as TriMesh::saveAsMesh writes it and read back."""
    import os
    src = os.path.join(os.path.dirname(__file__), "golden", "obj", "disc7.obj")
    V, UV0, F, FUV = capi.obj_read(src)
    assert FUV is not None and np.array_equal(FUV, F)
    ctx.set_mesh(V, F, fixed=(0,))
    UV, nl, nb, _ = ctx.tutte(precond=capi.PRECOND_AMG, rel_tol=1e-12, max_iter=1000)
    assert nl == 1 and nb == 24 and O.check_inversion(F, UV)
    ctx.trace_open(str(tmp_path))
    ctx.newton_init(a0=1.0, rel_gl2_tol=1e-12, allow_edec_rel_tol=True, precond=capi.PRECOND_AMG,
                    pcg_rel_tol=1e-8, pcg_max_iter=1000)
    for _ in range(50):
        s, done = ctx.newton_step()
        if done:
            break
    assert done
    ctx.trace_close()
    UVf = ctx.get_uv()
    feat = O.RestFeatures(V, F)
    assert O.check_inversion(F, UVf)
    assert O.energy(F, UVf, feat) < O.energy(F, UV, feat)
    out = str(tmp_path / "result_mesh.obj")
    capi.obj_write_mesh(out, V, UVf, F, F, scale_uv=True)
    V2, UV2, F2, FUV2 = capi.obj_read(out)
    np.testing.assert_allclose(V2, V, rtol=6e-7, atol=1e-12)
    np.testing.assert_array_equal(F2, F)
    span = (UVf.max(0) - UVf.min(0)).max()
    np.testing.assert_allclose(UV2, (UVf - UVf.min(0)) / span, rtol=0, atol=1e-6)
    assert len(_read(tmp_path / "gradientPerIter.txt")) >= 2
