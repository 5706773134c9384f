"""Writes tests/golden/ref_vectors.npz from the reference's own code.

Run in the container that holds the reference tree, after
``make -C oracle all ref``. The outputs come from the reference's
EigenLibSolver.cpp (set_pattern / update_a / SimplicialLDLT solve) and
IglUtils::makePD, built unmodified into oracle/_ref/libref.so. The inputs are
this project's synthetic meshes and the oracle's element Hessians (data, not
reference code).

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import oracle as O  # noqa: E402
from optcuts_amd.meshgen import height_field_disc  # noqa: E402


# Hand-written OBJ fixtures: TriMesh::save style; mixed syntax (groups,
# smoothing, materials, normals, i/t/n and negative indices, trailing
# comments, tabs, 3-value texture coordinates); positions and faces only.
OBJ_TEXTS = {
    "basic": """# two triangles, texture indices
v 0.000000e+00 0.000000e+00 0.000000e+00
v 1.000000e+00 0.000000e+00 0.000000e+00
v 0.000000e+00 1.000000e+00 2.500000e-01

v 1.000000e+00 1.000000e+00 -3.125000e-02
vt 0.000000e+00 0.000000e+00
vt 1.000000e+00 0.000000e+00
vt 0.000000e+00 1.000000e+00
vt 1.000000e+00 1.000000e+00
f 1/1 2/2 3/3
f 2/2 4/4 3/3
""",
    "mixed_syntax": """####
# exported by a tool
####
mtllib scene.mtl
o patch
g patch_group
v 0.1 0.2 0.3 # first
v\t1.5 -2.25   3e-3
v 4 5 6
v -1e2 2.5e+1 7
vn 0 0 1
vn 0 0 1
vn 0 1 0
vt 0.25 0.5 0
vt 0.75 0.5 0
vt 0.5 1.0 0
vt 0.0 0.125 0
usemtl skin
s off
f 1/1/1 2/2/2 3/3/3
f -3/-3/-2 -1/-1/-1 -2/-2/-3
s 1
f 4/4/1 1/1/2 3/3/3
""",
    "no_uv": """v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0.5 0.5 1
f 1 2 5
f 2//1 3//1 5//1
f 3 4 5
f 4 1 5
""",
}

def distorted_uv(V, F, UV, seed, amp):
    """A random smooth distortion that keeps every triangle CCW."""
    rng = np.random.default_rng(seed)
    for _ in range(50):
        k = rng.normal(size=(2, 2)) * amp
        UVd = UV + np.stack([k[0, 0] * np.sin(3 * UV[:, 1]) + k[0, 1] * UV[:, 0] ** 2,
                             k[1, 0] * np.cos(2 * UV[:, 0]) + k[1, 1] * UV[:, 1] ** 2], 1)
        if O.check_inversion(F, UVd):
            return UVd
        amp *= 0.5
    return UV.copy()


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    out = {}
    # --- CSR pattern, triplet accumulation and LDLT direction on three grids
    for m, seed in ((3, 11), (11, 12), (21, 13)):
        V, F, UV = height_field_disc(m, seed=seed)
        UVd = distorted_uv(V, F, UV, seed, 0.05)
        nV = V.shape[0]
        feat = O.RestFeatures(V, F)
        ptr, nbr = O.vertex_neighbors(nV, F)
        H = O.element_hessians(F, UVd, feat)
        I, J, Vt = O.hessian_triplets(F, H, fixed=(0,), scale=1.0)
        g = O.gradient(F, UVd, feat)
        S = O.RefSolver()
        ia, ja = S.set_pattern(nV, ptr, nbr, fixed=(0,))
        a = S.update_a(I, J, Vt)
        S.factorize()
        d = S.solve(-g)
        # a second right-hand side (a rotation-free random vector, fixed DOFs 0)
        rng = np.random.default_rng(seed + 100)
        b2 = rng.normal(size=2 * nV)
        b2[:2] = 0.0
        d2 = S.solve(b2)
        p = "grid%d_" % m
        out.update({p + "Vrest": V, p + "F": F, p + "UV": UVd, p + "ia": ia, p + "ja": ja, p + "a": a,
                    p + "g": g, p + "d": d, p + "b2": b2, p + "d2": d2})
    # --- makePD on random symmetric 6x6 and on element Hessians
    rng = np.random.default_rng(7)
    A = rng.normal(size=(100, 6, 6))
    A = 0.5 * (A + A.transpose(0, 2, 1))
    V, F, UV = height_field_disc(11, seed=5)
    UVd = distorted_uv(V, F, UV, 5, 0.3)
    feat = O.RestFeatures(V, F)
    He = O.element_hessians(F, UVd, feat, project=False)
    Hin = np.concatenate([A, He[:100]], 0)
    Hpd = np.stack([O.ref_make_pd6(h) for h in Hin])
    lam = np.stack([O.ref_eigvals6(h) for h in Hin])
    out.update({"pd_in": Hin, "pd_out": Hpd, "pd_eigvals": lam})
    # --- Tutte initial maps (main.cpp:1313-1364, 1486-1532): discs, a disc with a hole,
    # three charts in one mesh with shuffled triangles; layout 0 = longest loop per
    # component, layout 1 = every loop
    from optcuts_amd.meshgen import union_of_charts
    meshes = {}
    for name, (m, seed) in {"tutte11": (11, 12), "tutte21": (21, 13)}.items():
        meshes[name] = height_field_disc(m, seed=seed)[:2]
    V, F, _ = height_field_disc(21, seed=14)
    cen = V[F].mean(1)[:, :2]
    F = F[np.linalg.norm(cen - 0.5, axis=1) > 0.18]
    used = np.unique(F)
    remap = -np.ones(V.shape[0], np.int64)
    remap[used] = np.arange(len(used))
    meshes["tutte21h"] = (V[used], remap[F].astype(np.int32))
    V, F, _, _, _, _ = union_of_charts([height_field_disc(m, seed=s) for m, s in ((7, 15), (9, 16), (5, 17))])
    F = F[np.random.default_rng(18).permutation(F.shape[0])]
    meshes["tutte3c"] = (V, F)
    for name, (V, F) in meshes.items():
        out.update({name + "_Vrest": V, name + "_F": F})
        for layout in (0, 1):
            UVt, loops = O.ref_tutte(V, F, layout)
            key = "%s_l%d" % (name, layout)
            out.update({key + "_UV": UVt, key + "_bnd": np.concatenate(loops),
                        key + "_loops": np.array([len(l) for l in loops], np.int32)})
    # --- OBJ files (main.cpp:1153 reads them with igl::readOBJ): fixtures written
    # here, parsed by the vendored libigl reader
    obj_dir = os.path.join(HERE, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    texts = dict(OBJ_TEXTS)
    Vd, Fd, UVd = height_field_disc(7, seed=3)
    lines = ["v %e %e %e" % tuple(v) for v in Vd] + ["vt %e %e" % tuple(t) for t in UVd]
    lines += ["f %d/%d %d/%d %d/%d" % (a, a, b, b, c, c) for a, b, c in Fd + 1]
    texts["disc7"] = "\n".join(lines) + "\n"
    for name, text in sorted(texts.items()):
        path = os.path.join(obj_dir, name + ".obj")
        with open(path, "w") as f:
            f.write(text)
        V, TC, F, FTC = O.ref_read_obj(path)
        out.update({"obj_%s_V" % name: V, "obj_%s_TC" % name: TC, "obj_%s_F" % name: F, "obj_%s_FTC" % name: FTC})
    # --- std::ostream formatting of doubles (TriMesh::save: std::scientific; the
    # Optimizer traces: default flags)
    rng = np.random.default_rng(21)
    vals = np.concatenate([
        rng.standard_normal(60) * 10.0 ** rng.integers(-12, 12, 60),
        rng.uniform(-1, 1, 40),
        np.array([0.0, -0.0, 1.0, -1.0, 0.1, 1e-300, 5e-324, 1.7976931348623157e308, 123456789.0, 1e6, 999999.5,
                  1e-5, 0.0001, 4.812089839905548, 2.5e-7, 100000.0, 1234567.0, 9.9999995, 0.99999995]),
    ])
    fmt = {"values_hex": [float(v).hex() for v in vals],
           "scientific": [O.ref_stream_double(v, 1) for v in vals],
           "default": [O.ref_stream_double(v, 0) for v in vals]}
    with open(os.path.join(HERE, "stream_format.json"), "w") as f:
        json.dump(fmt, f, indent=0)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
