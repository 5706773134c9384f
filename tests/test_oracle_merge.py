"""The oracle's restatement of TriMesh::queryMerge's per-pair screen (TriMesh.cpp:801-908), pinned by hand-built
cases whose answers follow from the geometry (a feasible midpoint is kept exactly; one violated edge moves the
point along that edge's normal to eps inside it; half-planes that cannot hold together leave the pair infeasible)
and by properties on many random pairs (a projected point inverts no incident triangle; an infeasible one inverts
one; an unprojected one is the exact midpoint)."""
import numpy as np
import pytest

from oracle import oracle as O
from optcuts_amd.meshgen import height_field_disc

from _helpers import distort


def doubled_areas(P, F):
    a, b, c = P[F[:, 0]], P[F[:, 1]], P[F[:, 2]]
    return (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])


def slit_case(b2_guard):
    """Fork vertex f at the origin, two copies b1 (above the slit) and b2 (below) of a vertex at x = 1. b1 has a
    triangle whose opposite edge is the line y = 0.15 (b1 must stay above it); with b2_guard, b2 has one with the
    opposite edge y = -0.15 (b2 must stay below it)."""
    P = {"f": (0, 0), "b1": (1, 0.2), "b2": (1, -0.2), "p1": (0.5, 1), "p2": (0.5, -1), "s1": (0.5, 0.15),
         "r1": (1.5, 0.15), "s2": (0.5, -0.15), "r2": (1.5, -0.15)}
    names = list(P)
    idx = {n: i for i, n in enumerate(names)}
    tris = [("f", "b1", "p1"), ("b1", "s1", "r1"), ("f", "p2", "b2")]
    if b2_guard:
        tris.append(("b2", "r2", "s2"))
    UV = np.array([P[n] for n in names], np.float64)
    F = np.array([[idx[a], idx[b], idx[c]] for a, b, c in tris], np.int32)
    assert (doubled_areas(UV, F) > 0).all()
    row = np.array([[idx["f"], idx["b1"], idx["f"], idx["b2"]]], np.int32)
    return UV, F, row, idx


def test_feasible_midpoint_is_kept_exactly():
    UV, F, row, idx = slit_case(False)
    UV = UV.copy()
    UV[idx["b1"]] = (1.0, 0.3)  # midpoint (1, 0.05) is below y = 0.15: move b2 up so the midpoint clears it
    UV[idx["b2"]] = (1.0, 0.1)
    assert (doubled_areas(UV, F) > 0).all()
    st, mp, n = O.merge_screen(F, UV, row, 1.0)
    assert st[0] == 0 and n == 1
    assert mp[0, 0] == 1.0 and mp[0, 1] == (0.3 + 0.1) / 2.0


def test_one_violated_edge_projects_along_its_normal_to_eps_inside():
    UV, F, row, idx = slit_case(False)
    ael = 0.5
    eps = 1e-6 * ael
    st, mp, n = O.merge_screen(F, UV, row, ael)
    assert st[0] == 0 and n == 1
    # the midpoint (1, 0) is below the edge y = 0.15 of b1's second triangle; the edge's normal is vertical, so
    # x is unchanged and the point ends eps above the edge
    assert mp[0, 0] == 1.0
    assert mp[0, 1] == pytest.approx(0.15 + eps, abs=1e-15)
    Q = UV.copy()
    Q[idx["b1"]] = Q[idx["b2"]] = mp[0]
    assert (doubled_areas(Q, F) > 0).all()


def test_incompatible_half_planes_are_infeasible():
    UV, F, row, idx = slit_case(True)
    st, mp, n = O.merge_screen(F, UV, row, 0.5)
    assert st[0] == 2 and n == 0
    Q = UV.copy()
    Q[idx["b1"]] = Q[idx["b2"]] = mp[0]
    assert (doubled_areas(Q, F) < 0).any()


def test_rows_that_are_not_zipper_bottoms_are_skipped():
    UV, F, row, idx = slit_case(False)
    rows = np.array([[idx["f"], idx["b1"], idx["p1"], idx["b2"]], [idx["b1"], idx["f"], idx["b2"], idx["f"]]],
                    np.int32)
    st, mp, n = O.merge_screen(F, UV, rows, 0.5)
    # the second row forks at its second column: the merge vertices are b1 and b2 again
    assert st[0] == 1 and st[1] == 0 and n == 1


def random_rows(F, nV, n, seed):
    """Rows [f, b', f, b''] and [a', f, a'', f] of a vertex f and two of its neighbours, plus rows with four
    distinct vertices."""
    rng = np.random.default_rng(seed)
    nbr = [set() for _ in range(nV)]
    for t in F:
        for k in range(3):
            nbr[t[k]].add(int(t[(k + 1) % 3]))
            nbr[t[(k + 1) % 3]].add(int(t[k]))
    rows = []
    while len(rows) < n:
        f = int(rng.integers(nV))
        nb = sorted(nbr[f])
        if len(nb) < 2:
            continue
        a, b = rng.choice(nb, 2, replace=False)
        kind = rng.integers(3)
        if kind == 0:
            rows.append([f, a, f, b])
        elif kind == 1:
            rows.append([a, f, b, f])
        else:
            rows.append([f, a, int(rng.integers(nV)), b])
    return np.array(rows, np.int32)


def incident(F, v):
    return np.flatnonzero((F == v).any(axis=1))


@pytest.mark.parametrize("seed,amp", [(0, 0.0), (1, 0.1), (2, 0.15)])
def test_random_pairs_properties(seed, amp):
    V, F, UV = height_field_disc(21, seed=seed)
    if amp:
        UV = distort(UV, seed, amp=amp)
        assert O.check_inversion(F, UV)
    rows = random_rows(F, V.shape[0], 400, seed)
    ael = O.RestFeatures(V, F).avgEdgeLen
    st, mp, n = O.merge_screen(F, UV, rows, ael)
    assert n == int((st == 0).sum())
    counts = {0: 0, 1: 0, 2: 0, "projected": 0}
    for r, row in enumerate(rows):
        if row[0] == row[2]:
            va, vb = row[1], row[3]
        elif row[1] == row[3]:
            va, vb = row[0], row[2]
        else:
            assert st[r] == 1
            counts[1] += 1
            continue
        tris = np.union1d(incident(F, va), incident(F, vb))
        mid = (UV[va] + UV[vb]) / 2.0
        Q = UV.copy()
        Q[va] = Q[vb] = mid
        mid_ok = (doubled_areas(Q, F[tris]) >= 0).all()
        Q[va] = Q[vb] = mp[r]
        final_ok = (doubled_areas(Q, F[tris]) >= 0).all()
        counts[int(st[r])] += 1
        if mid_ok:
            assert st[r] == 0 and (mp[r] == mid).all()
        else:
            counts["projected"] += 1
            assert st[r] in (0, 2)
            assert final_ok == (st[r] == 0)
            if st[r] == 0:  # strictly inside every half-plane
                assert (doubled_areas(Q, F[tris]) > 0).all()
    print(seed, amp, counts)
    assert counts[0] > 0 and counts[1] > 0 and counts["projected"] > 0
