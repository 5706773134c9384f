"""The host set-up under sanitizers (CPU only): the Newton-system pattern, slot maps and multigrid structure
(oc_system_pattern.hpp, oc_amg_structure.hpp: from scratch, reusing rows after an air-mesh change, several builds
on several host threads at once) and the air mesh's Lawson re-mesh (oc_air_mesh.hpp), built as one executable
(tests/host/sanitize_main.cpp) with AddressSanitizer + UndefinedBehaviorSanitizer, and the concurrent builds
again with ThreadSanitizer. Any report fails the test (the runtimes abort on the first error)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from _helpers import flip_air_quads, grid_air_mesh
from oracle import oracle as O
from optcuts_amd import scaffold as SC
from optcuts_amd.meshgen import height_field_disc

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "sanitize_main.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def combined(nV, air):
    if air is None:
        return nV, np.zeros((0, 3), np.int32)
    airUV, Fair, bnd = air
    nb = len(bnd)
    l2g = np.concatenate([bnd, nV + np.arange(airUV.shape[0] - nb)]).astype(np.int32)
    return nV + airUV.shape[0] - nb, l2g[Fair]


def i32(a):
    return np.ascontiguousarray(a, np.int32).ravel().tobytes()


def reuse_case(m, air1, air2, fixed=(0,)):
    V, F, UV = height_field_disc(m, seed=0)
    nV = V.shape[0]
    nVt1, Fa1 = combined(nV, air1)
    nVt2, Fa2 = combined(nV, air2)
    h = [1, nV, F.shape[0], len(fixed), nVt1, Fa1.shape[0], nVt2, Fa2.shape[0]]
    return i32(h) + i32(F) + i32(fixed) + i32(Fa1) + i32(Fa2)


def concurrent_case(m, threads):
    V, F, UV = height_field_disc(m, seed=0)
    nV = V.shape[0]
    nVt, Fa = combined(nV, grid_air_mesh(m, rings=2))
    h = [2, nV, F.shape[0], 1, nVt, Fa.shape[0], threads, 0]
    return i32(h) + i32(F) + i32([0]) + i32(Fa)


def flips_case(seed):
    V, F, UV = height_field_disc(21, seed=4)
    ael = O.RestFeatures(V, F).avgEdgeLen
    airUV, Fair, bnd = SC.initial_air_mesh(UV, F, ael)
    nb = len(bnd)
    rng = np.random.default_rng(seed)
    move = rng.normal(size=(airUV.shape[0] - nb, 2)) * 0.3 * ael
    P = airUV
    for f in (1.0, 0.5, 0.25, 0.125, 0.0):
        Q = airUV.copy()
        Q[nb:] += f * move
        a, b, c = Q[Fair[:, 0]], Q[Fair[:, 1]], Q[Fair[:, 2]]
        if (((b - a)[:, 0] * (c - a)[:, 1] - (b - a)[:, 1] * (c - a)[:, 0]) > 0).all():
            P = Q
            break
    h = [3, P.shape[0], Fair.shape[0], 0, 0, 0, 0, 0]
    return i32(h) + np.ascontiguousarray(P, np.float64).tobytes() + i32(Fair)


def build_and_run(tmp_path, flags, cases, name):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread"] + flags +
                   [SRC, "-o", exe], check=True)
    data = tmp_path / (name + ".bin")
    data.write_bytes(b"".join(cases))
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:second_deadlock_stack=1"
    r = subprocess.run([exe, str(data)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert ("%d cases passed" % len(cases)) in r.stdout


def test_host_setup_under_address_and_undefined_sanitizers(tmp_path):
    m = 41
    air2 = grid_air_mesh(m, rings=2)
    cases = [
        reuse_case(m, air2, grid_air_mesh(m, rings=3)),
        reuse_case(61, grid_air_mesh(61, rings=2), flip_air_quads(grid_air_mesh(61, rings=2), 1)),
        reuse_case(m, None, air2),
        reuse_case(m, air2, None),
        reuse_case(m, air2, flip_air_quads(air2, 5), fixed=(0, m - 1, (m // 2) * m + m // 2)),
        concurrent_case(61, 4),
        flips_case(0),
        flips_case(1),
    ]
    build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-static-libasan"],
                  cases, "san_asan")


def test_concurrent_host_builds_under_thread_sanitizer(tmp_path):
    cases = [concurrent_case(61, 4), reuse_case(41, grid_air_mesh(41, rings=2), grid_air_mesh(41, rings=3))]
    build_and_run(tmp_path, ["-fsanitize=thread", "-static-libtsan"], cases, "san_tsan")
