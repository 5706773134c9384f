"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` import this module, and only as the checker. ``liboracle.so`` is
the project's own FP64 restatement of the reference hot path
(``oracle/sd_oracle.cpp``); ``libref.so`` wraps the reference's own
``EigenLibSolver.cpp`` and ``IglUtils::makePD`` compiled from the reference
tree (present only where that tree was available at build time).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build():
    """Compile the oracle (and the reference harness when the tree exists)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_rest_features.restype = C.c_int
        L.orc_rest_features.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_double] + [_dp] * 8 + [
            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_energy_per_elem.restype = None
        L.orc_energy_per_elem.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_energy.restype = C.c_double
        L.orc_energy.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int]
        L.orc_gradient.restype = None
        L.orc_gradient.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                   C.c_int, _ip, _dp]
        L.orc_local_gradients.restype = None
        L.orc_local_gradients.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_div_grad_per_vert.restype = None
        L.orc_div_grad_per_vert.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_split_candidates.restype = C.c_int
        L.orc_split_candidates.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, C.c_double, _ip,
                                           C.POINTER(C.c_int)]
        L.orc_merge_screen.restype = C.c_int
        L.orc_merge_screen.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_double, C.c_int, _ip, _ip, _dp]
        L.orc_make_pd.restype = None
        L.orc_make_pd.argtypes = [C.c_int, _dp]
        L.orc_eigh.restype = None
        L.orc_eigh.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_element_hessians.restype = None
        L.orc_element_hessians.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                           C.c_int, _dp]
        L.orc_hessian_triplets.restype = C.c_int
        L.orc_hessian_triplets.argtypes = [C.c_int, _ip, _dp, C.c_int, _ip, C.c_double, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
        L.orc_set_pattern.restype = C.c_int
        L.orc_set_pattern.argtypes = [C.c_int, _ip, _ip, C.c_int, _ip, C.c_void_p, C.c_void_p]
        L.orc_update_a.restype = C.c_int
        L.orc_update_a.argtypes = [C.c_int, _ip, _ip, C.c_int, _ip, _ip, _dp, _dp]
        L.orc_ldlt_factor.restype = C.c_void_p
        L.orc_ldlt_factor.argtypes = [C.c_int, _ip, _ip, _dp]
        L.orc_ldlt_solve.restype = None
        L.orc_ldlt_solve.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_ldlt_nnz.restype = C.c_long
        L.orc_ldlt_nnz.argtypes = [C.c_void_p]
        L.orc_ldlt_free.restype = None
        L.orc_ldlt_free.argtypes = [C.c_void_p]
        L.orc_init_step_size.restype = C.c_double
        L.orc_init_step_size.argtypes = [C.c_int, _ip, _dp, _dp, C.c_double]
        L.orc_check_inversion.restype = C.c_int
        L.orc_check_inversion.argtypes = [C.c_int, _ip, _dp]
        L.orc_seam_sparsity.restype = C.c_double
        L.orc_seam_sparsity.argtypes = [C.c_int, _ip, _dp, _dp, C.c_double, C.c_double, C.c_int]
        L.orc_newton_run.restype = C.c_int
        L.orc_newton_run.argtypes = [C.c_int, _dp, C.c_int, _ip, _dp, C.c_double, C.c_int, C.c_double,
                                     C.c_int, _dp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(_REF_PATH)


def ref():
    """The reference-code harness (EigenLibSolver + makePD), or None."""
    global _ref
    if _ref is None and ref_available():
        R = C.CDLL(_REF_PATH)
        R.ref_solver_create.restype = C.c_void_p
        R.ref_solver_destroy.argtypes = [C.c_void_p]
        R.ref_set_pattern.restype = C.c_int
        R.ref_set_pattern.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_int, _ip]
        R.ref_get_pattern.argtypes = [C.c_void_p, _ip, _ip]
        R.ref_update_a.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp]
        R.ref_get_a.argtypes = [C.c_void_p, _dp]
        R.ref_analyze_factorize.restype = C.c_int
        R.ref_analyze_factorize.argtypes = [C.c_void_p]
        R.ref_solve.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        R.ref_make_pd6.argtypes = [_dp]
        R.ref_eigvals6.argtypes = [_dp, _dp]
        R.ref_tutte.restype = C.c_int
        R.ref_tutte.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, _dp, _ip, _ip]
        R.ref_read_obj.restype = C.c_int
        R.ref_read_obj.argtypes = [C.c_char_p, _ip, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        R.ref_stream_double.restype = C.c_int
        R.ref_stream_double.argtypes = [C.c_double, C.c_int, C.c_char_p, C.c_int]
        _ref = R
    return _ref


# ----------------------------------------------------------------------------
# numpy-level helpers

def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class RestFeatures:
    """Per-triangle rest features (TriMesh.cpp:343-404)."""

    def __init__(self, Vrest, F, area_thres=0.0):
        Vrest = _f64(Vrest)
        F = _i32(F)
        nF = F.shape[0]
        arrs = [np.zeros(nF) for _ in range(8)]
        sa = C.c_double()
        ael = C.c_double()
        st = lib().orc_rest_features(Vrest.shape[0], Vrest, nF, F, float(area_thres), *arrs, C.byref(sa),
                                     C.byref(ael))
        if st != 0:
            raise ValueError("zero-area rest triangle")
        (self.triArea, self.triAreaSq, self.e0SqLen, self.e1SqLen, self.e0dote1, self.e0d, self.e1d,
         self.e01d) = arrs
        self.surfaceArea = sa.value
        self.avgEdgeLen = ael.value


def energy_per_elem(F, UV, feat, uniform=False):
    F = _i32(F)
    out = np.zeros(F.shape[0])
    lib().orc_energy_per_elem(F.shape[0], F, _f64(UV), feat.triArea, feat.triAreaSq, feat.e0SqLen, feat.e1SqLen,
                              feat.e0dote1, feat.surfaceArea, int(uniform), out)
    return out


def energy(F, UV, feat, uniform=False):
    F = _i32(F)
    return lib().orc_energy(F.shape[0], F, _f64(UV), feat.triArea, feat.triAreaSq, feat.e0SqLen, feat.e1SqLen,
                            feat.e0dote1, feat.surfaceArea, int(uniform))


def gradient(F, UV, feat, fixed=(0,), uniform=False):
    F = _i32(F)
    UV = _f64(UV)
    nV = UV.shape[0]
    g = np.zeros(2 * nV)
    fx = _i32(np.array(sorted(fixed), dtype=np.int32))
    lib().orc_gradient(nV, F.shape[0], F, UV, feat.triArea, feat.triAreaSq, feat.e0SqLen, feat.e1SqLen,
                       feat.e0dote1, feat.surfaceArea, int(uniform), len(fx), fx, g)
    return g


def local_gradients(F, UV, feat):
    """SymDirichletEnergy::computeLocalGradient (SymDirichletEnergy.cpp:215-256): (nF, 3, 2)."""
    F = _i32(F)
    out = np.zeros((F.shape[0], 3, 2))
    lib().orc_local_gradients(F.shape[0], F, _f64(UV), feat.triArea, feat.triAreaSq, feat.e0SqLen, feat.e1SqLen,
                              feat.e0dote1, feat.surfaceArea, out)
    return out


def div_grad_per_vert(F, UV, feat):
    """SymDirichletEnergy::computeDivGradPerVert (SymDirichletEnergy.cpp:108-154): (nV,)."""
    F = _i32(F)
    UV = _f64(UV)
    out = np.zeros(UV.shape[0])
    lib().orc_div_grad_per_vert(UV.shape[0], F.shape[0], F, UV, feat.triArea, feat.triAreaSq, feat.e0SqLen,
                                feat.e1SqLen, feat.e0dote1, feat.surfaceArea, out)
    return out


def split_candidates(F, nV, div_grad, vert_weight=None, split_interior=False, filter_exp_in=0.6):
    """The candidate filter of TriMesh::querySplit (TriMesh.cpp:556-633, propagate = false):
    (candidates in key order, size of the sorted candidate map)."""
    F = _i32(F)
    w = np.ones(nV) if vert_weight is None else _f64(vert_weight)
    cand = np.zeros(max(nV, 2), np.int32)
    ns = C.c_int(0)
    n = lib().orc_split_candidates(nV, F.shape[0], F, _f64(div_grad), w, int(split_interior), float(filter_exp_in),
                                   cand, C.byref(ns))
    return cand[:n].copy(), ns.value


def merge_screen(F, UV, cohE, avg_edge_len):
    """The per-pair screen of TriMesh::queryMerge (TriMesh.cpp:801-905, propagate = false) before its local
    evaluation: (status per cohesive row: 0 candidate, 1 not a zipper bottom, 2 infeasible; merged positions
    (nCoh, 2); number of candidates)."""
    F = _i32(F)
    UV = _f64(UV)
    cohE = _i32(cohE).reshape(-1, 4)
    st = np.zeros(cohE.shape[0], np.int32)
    mp = np.zeros((cohE.shape[0], 2))
    n = lib().orc_merge_screen(UV.shape[0], F.shape[0], F, UV, float(avg_edge_len), cohE.shape[0], cohE, st,
                               mp.reshape(-1))
    return st, mp, n


def element_hessians(F, UV, feat, uniform=False, project=True):
    F = _i32(F)
    H = np.zeros((F.shape[0], 6, 6))
    lib().orc_element_hessians(F.shape[0], F, _f64(UV), feat.triArea, feat.triAreaSq, feat.e0d, feat.e1d,
                               feat.e01d, feat.surfaceArea, int(uniform), int(project), H.reshape(-1))
    return H


def make_pd(H):
    H = _f64(H).copy()
    n = H.shape[0]
    lib().orc_make_pd(n, H.reshape(-1))
    return H


def eigh(A):
    A = _f64(A)
    n = A.shape[0]
    lam = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_eigh(n, A.reshape(-1), lam, V.reshape(-1))
    return lam, V


def hessian_triplets(F, Helem, fixed=(0,), scale=1.0):
    F = _i32(F)
    He = _f64(Helem).reshape(-1)
    fx = _i32(np.array(sorted(fixed), dtype=np.int32))
    n = lib().orc_hessian_triplets(F.shape[0], F, He, len(fx), fx, float(scale), None, None, None)
    I = np.zeros(n, np.int32)
    J = np.zeros(n, np.int32)
    V = np.zeros(n)
    lib().orc_hessian_triplets(F.shape[0], F, He, len(fx), fx, float(scale), I.ctypes.data, J.ctypes.data,
                               V.ctypes.data)
    return I, J, V


def vertex_neighbors(nV, F):
    """vNeighbor as CSR (TriMesh.cpp:436-447)."""
    F = np.asarray(F, dtype=np.int64)
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    key = np.unique(e[:, 0] * nV + e[:, 1])
    a, b = key // nV, key % nV
    ptr = np.zeros(nV + 1, np.int32)
    np.add.at(ptr, a + 1, 1)
    ptr = np.cumsum(ptr).astype(np.int32)
    return ptr, b.astype(np.int32)


def set_pattern(nV, nbr_ptr, nbr, fixed=(0,)):
    fx = _i32(np.array(sorted(fixed), dtype=np.int32))
    nbr_ptr, nbr = _i32(nbr_ptr), _i32(nbr)
    nnz = lib().orc_set_pattern(nV, nbr_ptr, nbr, len(fx), fx, None, None)
    ia = np.zeros(2 * nV + 1, np.int32)
    ja = np.zeros(nnz, np.int32)
    lib().orc_set_pattern(nV, nbr_ptr, nbr, len(fx), fx, ia.ctypes.data, ja.ctypes.data)
    return ia, ja


def update_a(ia, ja, I, J, V):
    a = np.zeros(len(ja))
    st = lib().orc_update_a(len(ia) - 1, _i32(ia), _i32(ja), len(I), _i32(I), _i32(J), _f64(V), a)
    if st != 0:
        raise ValueError("triplet outside the pattern")
    return a


class LDLT:
    """Own simplicial LDL^T on the upper CSR (restates EigenLibSolver's role)."""

    def __init__(self, ia, ja, a):
        self._ia, self._ja, self._a = _i32(ia), _i32(ja), _f64(a)
        self.n = len(ia) - 1
        self.h = lib().orc_ldlt_factor(self.n, self._ia, self._ja, self._a)
        if not self.h:
            raise np.linalg.LinAlgError("LDLT: zero pivot")

    def solve(self, b):
        x = np.zeros(self.n)
        lib().orc_ldlt_solve(self.h, _f64(b), x)
        return x

    def nnz(self):
        return lib().orc_ldlt_nnz(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ldlt_free(self.h)
            self.h = None


def init_step_size(F, UV, d, step=1.0):
    F = _i32(F)
    return lib().orc_init_step_size(F.shape[0], F, _f64(UV), _f64(d), float(step))


def check_inversion(F, UV):
    F = _i32(F)
    return bool(lib().orc_check_inversion(F.shape[0], F, _f64(UV)))


def seam_sparsity(cohE, Vrest, UV, avgEdgeLen, initSeamLen=0.0, triSoup=False):
    cohE = _i32(cohE).reshape(-1, 4)
    return lib().orc_seam_sparsity(cohE.shape[0], cohE, _f64(Vrest), _f64(UV), float(avgEdgeLen),
                                   float(initSeamLen), int(triSoup))


def standard_stretch(Vrest, F, UV):
    """TriMesh::computeStandardStretch (TriMesh.cpp:1585-1652): (stretch_l2, stretch_inf, stretch_shear,
    compress_inf) of the map, each triangle's deformation gradient from IglUtils::computeDeformationGradient
    (IglUtils.cpp:172-194: the rest triangle laid flat with its first edge on the x axis, the UV triangle likewise,
    F = X U^-1), L2 and shear averaged by rest area and square-rooted, stretches scaled by sqrt(UV area / rest area)."""
    P = np.asarray(Vrest, np.float64)[F]
    Q = np.asarray(UV, np.float64)[F]
    e0, e1 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    cr = np.cross(e0, e1)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    x1 = np.linalg.norm(e0, axis=1)
    x2 = np.stack([(e0 * e1).sum(1) / x1, np.linalg.norm(cr, axis=1) / x1], 1)
    u01, u02 = Q[:, 1] - Q[:, 0], Q[:, 2] - Q[:, 0]
    l01 = np.linalg.norm(u01, axis=1)
    U = np.zeros((len(F), 2, 2))
    U[:, 0, 0] = l01
    U[:, 0, 1] = (u01 * u02).sum(1) / l01
    U[:, 1, 1] = (u01[:, 0] * u02[:, 1] - u01[:, 1] * u02[:, 0]) / l01
    X = np.zeros((len(F), 2, 2))
    X[:, 0, 0] = x1
    X[:, :, 1] = x2
    G = X @ np.linalg.inv(U)
    a = (G[:, :, 0] ** 2).sum(1)
    b = (G[:, :, 0] * G[:, :, 1]).sum(1)
    c = (G[:, :, 1] ** 2).sum(1)
    t0 = a + c
    t1 = np.sqrt((a - c) ** 2 + 4.0 * b * b)
    tau = np.sqrt((t0 + t1) / 2.0)
    with np.errstate(invalid="ignore"):
        gamma = np.sqrt((t0 - t1) / 2.0)  # NaN when rounding makes t0 < t1: the reference's '<' skips it
    sa = area.sum()
    scale = np.sqrt(0.5 * (u01[:, 0] * u02[:, 1] - u01[:, 1] * u02[:, 0]).sum() / sa)
    return (np.sqrt((t0 / 2.0 * area).sum() / sa) * scale, tau.max() * scale,
            np.sqrt((b * b / a / c * area).sum() / sa), np.nanmin(gamma) * scale)


def newton_run(Vrest, F, UV, a0=1.0, max_iter=100, rel_gl2_tol=1e-12, allow_edec_rel_tol=True):
    """Plain Newton + own LDLT, no scaffold. Returns (trace rows, final UV, exit reason)."""
    UV = _f64(UV).copy()
    F = _i32(F)
    Vrest = _f64(Vrest)
    trace = np.zeros((max_iter + 1, 6))
    why = C.c_int()
    n = lib().orc_newton_run(Vrest.shape[0], Vrest, F.shape[0], F, UV.reshape(-1), float(a0), int(max_iter),
                             float(rel_gl2_tol), int(allow_edec_rel_tol), trace.reshape(-1), C.byref(why))
    if n < 0:
        raise RuntimeError("oracle newton_run failed with status %d" % n)
    return trace[: n + 1], UV, why.value


# ----------------------------------------------------------------------------
# reference harness helpers

class RefSolver:
    """The reference EigenLibSolver (set_pattern/update_a/factorize/solve)."""

    def __init__(self):
        self.R = ref()
        if self.R is None:
            raise RuntimeError("reference harness not built")
        self.h = self.R.ref_solver_create()

    def set_pattern(self, nV, nbr_ptr, nbr, fixed=(0,)):
        fx = _i32(np.array(sorted(fixed), dtype=np.int32))
        self.nnz = self.R.ref_set_pattern(self.h, nV, _i32(nbr_ptr), _i32(nbr), len(fx), fx)
        self.n = 2 * nV
        ia = np.zeros(self.n + 1, np.int32)
        ja = np.zeros(self.nnz, np.int32)
        self.R.ref_get_pattern(self.h, ia, ja)
        return ia, ja

    def update_a(self, I, J, V):
        self.R.ref_update_a(self.h, len(I), _i32(I), _i32(J), _f64(V))
        a = np.zeros(self.nnz)
        self.R.ref_get_a(self.h, a)
        return a

    def factorize(self):
        if self.R.ref_analyze_factorize(self.h) != 0:
            raise np.linalg.LinAlgError("reference factorize failed")

    def solve(self, b):
        x = np.zeros(self.n)
        self.R.ref_solve(self.h, self.n, _f64(b), x)
        return x

    def __del__(self):
        if getattr(self, "h", None):
            self.R.ref_solver_destroy(self.h)
            self.h = None


def ref_make_pd6(H):
    H = _f64(H).copy()
    ref().ref_make_pd6(H.reshape(-1))
    return H


def ref_eigvals6(H):
    lam = np.zeros(6)
    ref().ref_eigvals6(_f64(H).reshape(-1), lam)
    return lam


def ref_tutte(V, F, layout=0):
    """The reference's Tutte initial map through the vendored libigl
    (main.cpp:1313-1364 for layout 1, 1380-1395 + 1486-1532 for layout 0).
    Returns (UV, list of boundary loops)."""
    V, F = _f64(V), _i32(F)
    UV = np.zeros((V.shape[0], 2))
    bnd = np.zeros(V.shape[0] + 1, np.int32)
    lens = np.zeros(V.shape[0] + 1, np.int32)
    n = ref().ref_tutte(V.shape[0], V.reshape(-1), F.shape[0], F.reshape(-1), int(layout), UV.reshape(-1),
                        bnd, lens)
    if n < 0:
        raise ValueError("a component has no boundary")
    cut = np.concatenate([[0], np.cumsum(lens[:n])])
    return UV, [bnd[cut[k]:cut[k + 1]].copy() for k in range(n)]


def _edge_counts(F):
    count = {}
    for f in range(F.shape[0]):
        for i in range(3):
            a, b = int(F[f, i]), int(F[f, (i + 1) % 3])
            key = (min(a, b), max(a, b))
            count[key] = count.get(key, 0) + 1
    return count


def facet_components(F):
    """Edge-connected triangle components (igl/facet_components.cpp), numbered in
    order of their lowest triangle index."""
    F = np.asarray(F)
    owners = {}
    for f in range(F.shape[0]):
        for i in range(3):
            a, b = int(F[f, i]), int(F[f, (i + 1) % 3])
            owners.setdefault((min(a, b), max(a, b)), []).append(f)
    C = -np.ones(F.shape[0], np.int64)
    nc = 0
    for g in range(F.shape[0]):
        if C[g] >= 0:
            continue
        C[g] = nc
        stack = [g]
        while stack:
            f = stack.pop()
            for i in range(3):
                a, b = int(F[f, i]), int(F[f, (i + 1) % 3])
                for h in owners[(min(a, b), max(a, b))]:
                    if C[h] < 0:
                        C[h] = nc
                        stack.append(h)
        nc += 1
    return C


def boundary_loops(F, nV):
    """Every boundary loop as libigl's boundary_loop walks them (boundary_loop.cpp,
    triangle_triangle_adjacency.cpp, is_border_vertex.cpp): boundary edges are
    triangle edges (F[f,i], F[f,i+1]) whose undirected edge no other triangle
    has; each loop starts at the smallest unvisited border vertex and moves to
    the next vertex of the first incident triangle (ascending) whose boundary
    edge leaves the current one to an unvisited vertex."""
    F = np.asarray(F)
    count = _edge_counts(F)
    bedge = np.zeros((F.shape[0], 3), bool)
    border = np.zeros(nV, bool)
    for f in range(F.shape[0]):
        for i in range(3):
            a, b = int(F[f, i]), int(F[f, (i + 1) % 3])
            if count[(min(a, b), max(a, b))] == 1:
                bedge[f, i] = True
                border[a] = border[b] = True
    vf = [[] for _ in range(nV)]
    for f in range(F.shape[0]):
        for i in range(3):
            vf[int(F[f, i])].append(f)
    unvisited = border.copy()
    loops = []
    for start in range(nV):
        if not unvisited[start]:
            continue
        loop = [start]
        unvisited[start] = False
        while True:
            v = loop[-1]
            nxt = -1
            for f in vf[v]:
                if not bedge[f].any():
                    continue
                loc = int(np.nonzero(F[f] == v)[0][0])
                vn = int(F[f, (loc + 1) % 3])
                if unvisited[vn] and bedge[f, loc]:
                    nxt = vn
                    break
            if nxt < 0:
                break
            loop.append(nxt)
            unvisited[nxt] = False
        loops.append(np.array(loop, np.int32))
    return loops


def boundary_loop(F, nV):
    """The first longest boundary loop (igl::boundary_loop's single-loop form)."""
    best = np.zeros(0, np.int32)
    for l in boundary_loops(F, nV):
        if len(l) > len(best):
            best = l
    return best


def tutte_loops(F, nV, layout=0):
    """Loops pinned by the Tutte map: layout 0 = the first longest loop of each
    facet component, walked over that component's triangles (main.cpp:1486-1512);
    layout 1 = every loop of the mesh (main.cpp:1313-1317)."""
    F = np.asarray(F)
    if layout == 1:
        loops = boundary_loops(F, nV)
        if not loops:
            raise ValueError("a component has no boundary")
        return loops
    C = facet_components(F)
    loops = []
    for k in range(int(C.max()) + 1):
        l = boundary_loop(F[C == k], nV)
        if len(l) == 0:
            raise ValueError("a component has no boundary")
        loops.append(l)
    return loops


def tutte_uv(V, F, layout=0):
    """Tutte initial map (main.cpp:1313-1364, 1486-1532): loop k on the unit circle
    by arc length of the rest positions (map_vertices_to_circle.cpp), shifted by
    (k % d, k // d) * 2.1 with d the smallest integer with d*d >= number of loops;
    the interior harmonic for the uniform Laplacian of IglUtils.cpp:39-60
    (igl::harmonic with k = 1: Q = -L, Q_II W_I = -Q_IB bc), solved directly.
    Returns (UV, loops)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    V, F = _f64(V), np.asarray(F)
    nV = V.shape[0]
    loops = tutte_loops(F, nV, layout)
    d = 1
    while d * d < len(loops):
        d += 1
    bnd = np.concatenate(loops)
    bcs = []
    for k, l in enumerate(loops):
        seg = np.linalg.norm(V[l[1:]] - V[l[:-1]], axis=1)
        length = np.concatenate([[0.0], np.cumsum(seg)])
        total = length[-1] + np.linalg.norm(V[l[0]] - V[l[-1]])
        ang = length * 2.0 * np.pi / total
        bcs.append(np.stack([np.cos(ang) + (k % d) * 2.1, np.sin(ang) + (k // d) * 2.1], 1))
    bc = np.concatenate(bcs)
    r, c, v = [], [], []
    for (a, b) in ((0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2)):
        r.append(F[:, a]); c.append(F[:, b]); v.append(np.ones(F.shape[0]))
    for a in range(3):
        r.append(F[:, a]); c.append(F[:, a]); v.append(-2.0 * np.ones(F.shape[0]))
    L = sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(nV, nV))
    Q = (-L).tocsr()
    isb = np.zeros(nV, bool)
    isb[bnd] = True
    I = np.nonzero(~isb)[0]
    W = np.zeros((nV, 2))
    W[bnd] = bc
    if len(I):
        rhs = -(Q[I][:, bnd] @ bc)
        W[I] = spla.spsolve(Q[I][:, I].tocsc(), rhs).reshape(-1, 2)
    return W, loops


def ref_read_obj(path):
    """igl::readOBJ(path, V, TC, N, F, FTC, FN) as main.cpp:1153 calls it (vendored libigl).
    Returns (V, TC, F, FTC) or None when the reader fails."""
    R = ref()
    sizes = np.zeros(7, np.int32)
    p = os.fsencode(path)
    if not R.ref_read_obj(p, sizes, None, None, None, None):
        return None
    V = np.zeros((sizes[0], sizes[1]))
    TC = np.zeros((sizes[2], sizes[3]))
    F = np.zeros((sizes[4], sizes[5]), np.int32)
    FTC = np.zeros((sizes[6], sizes[5]), np.int32)
    R.ref_read_obj(p, sizes, V.ctypes.data, TC.ctypes.data, F.ctypes.data, FTC.ctypes.data)
    return V, TC, F, FTC


def ref_stream_double(x, scientific):
    """x as std::ostream writes it, default flags or std::scientific."""
    buf = C.create_string_buffer(64)
    ref().ref_stream_double(float(x), int(scientific), buf, 64)
    return buf.value.decode()
