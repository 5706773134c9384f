"""Host set-up reuse after an air-mesh change (CPU only).

When only the air mesh changes, the Newton-system pattern, its slot maps and
the multigrid structure copy the rows of mesh vertices away from the air
triangles from the previous build. The copies must equal a build from scratch
array for array; tests/host/reuse_check.cpp builds both and names the first
array that differs."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from _helpers import flip_air_quads, grid_air_mesh
from optcuts_amd.meshgen import height_field_disc

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("reuse") / "reuse_check.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-pthread",
                    os.path.join(HERE, "host", "reuse_check.cpp"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    lib.reuse_check.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ctypes.c_int, ip,
                                ctypes.c_int, ctypes.c_int, ip, ip, ctypes.c_char_p, ctypes.c_int]
    lib.reuse_check.restype = ctypes.c_int
    lib.concurrent_check.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int, ctypes.c_int,
                                     ip, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.concurrent_check.restype = ctypes.c_int
    return lib


def combined(nV, air):
    """Air mesh (airUV, Fair local, bnd) -> (nVt, Fair in combined ids)."""
    if air is None:
        return nV, np.zeros((0, 3), np.int32)
    airUV, Fair, bnd = air
    nb = len(bnd)
    l2g = np.concatenate([bnd, nV + np.arange(airUV.shape[0] - nb)]).astype(np.int32)
    return nV + airUV.shape[0] - nb, l2g[Fair]


def renumber_interior(air, seed):
    """Same air triangles, interior air vertices in another order."""
    airUV, Fair, bnd = air
    nb = len(bnd)
    rng = np.random.default_rng(seed)
    perm = np.concatenate([np.arange(nb), nb + rng.permutation(airUV.shape[0] - nb)])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return airUV[perm], inv[Fair].astype(np.int32), bnd


def run(checker, m, air1, air2, fixed=(0,)):
    V, F, UV = height_field_disc(m, seed=0)
    nV = V.shape[0]
    nVt1, Fa1 = combined(nV, air1)
    nVt2, Fa2 = combined(nV, air2)
    stats = np.zeros(4, np.int32)
    what = ctypes.create_string_buffer(256)
    fx = np.asarray(fixed, np.int32)
    rc = checker.reuse_check(nV, F.shape[0], np.ascontiguousarray(F, np.int32), fx.size, fx, nVt1, Fa1.shape[0],
                             np.ascontiguousarray(Fa1.ravel()), nVt2, Fa2.shape[0],
                             np.ascontiguousarray(Fa2.ravel()), stats, what, 256)
    assert rc == 0, what.value.decode()
    return stats


@pytest.mark.parametrize("m", [41, 101])
def test_wider_air_band(checker, m):
    stats = run(checker, m, grid_air_mesh(m, rings=2), grid_air_mesh(m, rings=3))
    assert stats[0] >= 3  # a hierarchy with transfers was compared
    assert stats[1] == m * m - 4 * (m - 1)  # every interior mesh vertex kept


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_new_air_triangulation(checker, seed):
    m = 61
    air = grid_air_mesh(m, rings=2)
    stats = run(checker, m, air, flip_air_quads(air, seed))
    assert stats[0] >= 3 and stats[1] > 0


def test_renumbered_air_vertices(checker):
    m = 61
    air = grid_air_mesh(m, rings=3)
    run(checker, m, air, renumber_interior(air, 7))


def test_same_air_mesh(checker):
    m = 41
    air = grid_air_mesh(m, rings=2)
    run(checker, m, air, air)


def test_air_mesh_added_and_removed(checker):
    m = 41
    air = grid_air_mesh(m, rings=2)
    stats = run(checker, m, None, air)
    assert stats[1] == m * m - 4 * (m - 1)
    stats = run(checker, m, air, None)
    assert stats[1] == m * m - 4 * (m - 1)


def test_fixed_vertices_inside_and_on_the_boundary(checker):
    m = 41
    air = grid_air_mesh(m, rings=2)
    centre = (m // 2) * m + m // 2
    run(checker, m, air, flip_air_quads(air, 5), fixed=(0, m - 1, centre))


def test_concurrent_builds(checker):
    """Set-ups of several contexts on several host threads at once (one holds the worker pool, the others
    start their own threads) give the same structure as one set-up alone."""
    m = 101
    V, F, UV = height_field_disc(m, seed=0)
    nV = V.shape[0]
    nVt, Fa = combined(nV, grid_air_mesh(m, rings=2))
    what = ctypes.create_string_buffer(256)
    fx = np.asarray((0,), np.int32)
    rc = checker.concurrent_check(nV, F.shape[0], np.ascontiguousarray(F, np.int32), 1, fx, nVt, Fa.shape[0],
                                  np.ascontiguousarray(Fa.ravel()), 4, what, 256)
    assert rc == 0, what.value.decode()
