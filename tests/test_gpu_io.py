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


def _fields(line):
    return [float(x) for x in line.split()]


@pytest.mark.parametrize("fracture", [False, True])
def test_energy_trace_seam_column_matches_oracle(ctx, tmp_path, fracture):
    """A cut closed surface (icosphere, genus 0) with its cohesive edges: every energy line recomputes
    E_se = computeSeamSparsity(!fractureMode) / virtualRadius at the current UV, and in fracture mode the first
    column is lastEnergyVal + (1 - a0) E_se (Optimizer.cpp:706-728); checked against the oracle's seam length at the
    UV of each line, both modes."""
    from optcuts_amd import surfaces as S
    Vc, Fc, cohE, meta = S.workload("icosphere_nu10")
    ctx.set_mesh(Vc, Fc, fixed=(0,))
    ctx.tutte(capi.TUTTE_COMPONENTS, precond=capi.PRECOND_JACOBI, rel_tol=1e-13)
    feat = O.RestFeatures(Vc, Fc)
    vr = np.sqrt(feat.surfaceArea / np.pi)
    a0, init = 0.25, 0.125
    ctx.trace_open(str(tmp_path))
    ctx.trace_set_seams(cohE, init_seam_len=init, fracture_mode=fracture)
    E0 = ctx.newton_init(a0=a0, rel_gl2_tol=1e-12, allow_edec_rel_tol=False, pcg_rel_tol=1e-12,
                         pcg_max_iter=20000)
    states = [(E0, ctx.energy(), ctx.get_uv())]
    for k in range(4):
        s, done = ctx.newton_step()
        assert not s["stopped"]
        states.append((s["E_total"], s["E_SD"], ctx.get_uv()))
    ctx.trace_close()
    lines = _read(str(tmp_path / "energyValPerIter.txt"))
    assert len(lines) == len(states)
    for line, (E, esd, uv) in zip(lines, states):
        se = O.seam_sparsity(cohE, Vc, uv, feat.avgEdgeLen, init, triSoup=not fracture) / vr
        col0 = E + (1.0 - a0) * se if fracture else E
        got = _fields(line)
        assert len(got) == 4 and line == " ".join("%g" % x for x in got)
        np.testing.assert_allclose(got, [col0, esd, se, a0], rtol=2e-5)
    # the seams open up as the map relaxes, so the tri-soup (non-fracture) count is not constant
    if not fracture:
        assert len({ln.split()[2] for ln in lines}) >= 1


def test_info_file_matches_oracle(ctx, tmp_path):
    """info.txt (main.cpp:346-387) after a few steps: the caller's counters and timings, E_SD, seam length over
    virtualRadius, the standard stretch metrics and the initial seams printed as Eigen prints an integer matrix."""
    from optcuts_amd import surfaces as S
    Vc, Fc, cohE, meta = S.workload("icosphere_nu10")
    ctx.set_mesh(Vc, Fc, fixed=(0,))
    ctx.tutte(capi.TUTTE_COMPONENTS, precond=capi.PRECOND_JACOBI, rel_tol=1e-13)
    ctx.trace_set_seams(cohE, fracture_mode=True)
    ctx.newton_init(a0=0.001, rel_gl2_tol=1e-12, allow_edec_rel_tol=False, pcg_rel_tol=1e-10, pcg_max_iter=20000)
    for _ in range(3):
        ctx.newton_step()
    timings = np.arange(15) * 0.5 + 0.25
    p = str(tmp_path / "info.txt")
    ctx.info_write(p, meta["vertices"], meta["faces"], 3, 0, 0.999, timings, cohE)
    lines = open(p).read().split("\n")
    feat = O.RestFeatures(Vc, Fc)
    uv = ctx.get_uv()
    assert lines[0] == "%d %d" % (meta["vertices"], meta["faces"])
    assert lines[1] == "3 0 0 0 %g %g" % (0.999, 1.0 - 0.001)
    lbl = ["topo", "desc", "scaf", "enUp", "mtrComp", "mtrAssem", "symFac", "numFac", "backSolve", "lineSearch",
           "bSplit", "iSplit", "cMerge"]
    assert lines[2] == "0.0 0.0 %g %g " % (timings[0], timings[1]) + " ".join(
        "%s%g" % (l, t) for l, t in zip(lbl, timings[2:]))
    d, sl = _fields(lines[3])
    assert d == pytest.approx(O.energy(Fc, uv, feat), rel=1e-5)
    assert sl == pytest.approx(O.seam_sparsity(cohE, Vc, uv, feat.avgEdgeLen, 0.0, triSoup=False) /
                               np.sqrt(feat.surfaceArea / np.pi), rel=1e-5)
    np.testing.assert_allclose(_fields(lines[4]), O.standard_stretch(Vc, Fc, uv), rtol=1e-5)
    assert lines[5] == "initialSeams %d" % cohE.shape[0]
    w = max(len(str(int(x))) for x in cohE.ravel())
    for r in range(cohE.shape[0]):
        assert lines[6 + r] == " ".join(str(int(x)).rjust(w) for x in cohE[r])
    assert lines[6 + cohE.shape[0]] == "" and len(lines) == 7 + cohE.shape[0]
