"""ctypes binding of liboptcuts_gpu.so (the C ABI in include/optcuts_gpu.h).

The library is the product path: every call below runs HIP kernels on the
GPU. There is no CPU fallback; loading fails loudly when the library is not
built or no gfx950 device is present.
"""
import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liboptcuts_gpu.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "optcuts_gpu.h")

OC_OK = 0
STATUS = {
    0: "OC_OK", -1: "OC_ERR_ARG", -2: "OC_ERR_HIP", -3: "OC_ERR_STATE", -4: "OC_ERR_PATTERN",
    -5: "OC_ERR_INVERTED", -6: "OC_ERR_DEGENERATE", -7: "OC_ERR_NOT_CONVERGED", -8: "OC_ERR_BREAKDOWN",
    -9: "OC_ERR_NO_DEVICE",
}
PRECOND_JACOBI = 0
PRECOND_BLOCK_JACOBI = 1
PRECOND_AMG = 2
PRECOND_LEARNED = 3
PRECOND_IC0 = 4
PRECISION_FP64 = 0
PRECISION_MIXED = 1
TUTTE_COMPONENTS = 0
TUTTE_ALL_LOOPS = 1
SCAFFOLD_REFRESH = 0
SCAFFOLD_REMESH_WHEN_NEEDED = 1
SCAFFOLD_REMESH = 2


class OCError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class IterStats(C.Structure):
    _fields_ = [("E_total", C.c_double), ("E_SD", C.c_double), ("E_scaf", C.c_double), ("grad_sq", C.c_double),
                ("step", C.c_double), ("halvings", C.c_int), ("stopped", C.c_int), ("converged", C.c_int),
                ("pcg_iters", C.c_int), ("pcg_rel_res", C.c_double), ("pcg_fallback", C.c_int),
                ("pcg_iters_first", C.c_int), ("precond_setups", C.c_int), ("ir_passes", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_ctx = C.c_void_p

_SIGS = {
    "oc_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "oc_ctx_destroy": (C.c_int, [_ctx]),
    "oc_last_error": (C.c_char_p, [_ctx]),
    "oc_device_arch": (C.c_int, [_ctx, C.c_char_p, C.c_int]),
    "oc_set_mesh": (C.c_int, [_ctx, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, C.c_double]),
    "oc_set_uv": (C.c_int, [_ctx, _dp]),
    "oc_get_uv": (C.c_int, [_ctx, _dp]),
    "oc_mesh_stats": (C.c_int, [_ctx, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "oc_energy": (C.c_int, [_ctx, C.c_int, C.POINTER(C.c_double)]),
    "oc_energy_per_elem": (C.c_int, [_ctx, C.c_int, _dp]),
    "oc_energy_noalloc": (C.c_int, [_ctx, C.c_int, C.POINTER(C.c_double)]),
    "oc_gradient": (C.c_int, [_ctx, C.c_int, _dp]),
    "oc_hessian_triplets": (C.c_int, [_ctx, C.c_int, C.c_longlong, _vp, _vp, _vp, C.POINTER(C.c_longlong)]),
    "oc_element_terms": (C.c_int, [_ctx, C.c_int, _vp, _vp, _vp]),
    "oc_init_step_size": (C.c_int, [_ctx, _dp, C.POINTER(C.c_double)]),
    "oc_check_inversion": (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    "oc_div_grad_per_vert": (C.c_int, [_ctx, _dp]),
    "oc_pcg_graphs": (C.c_int, [_ctx, C.c_int]),
    "oc_seam_sparsity": (C.c_int, [_ctx, C.c_int, _ip, C.c_double, C.c_int, C.POINTER(C.c_double)]),
    "oc_split_candidates": (C.c_int, [_ctx, C.c_int, C.c_double, _vp, C.c_int, _ip, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "oc_merge_candidates": (C.c_int, [_ctx, C.c_int, _ip, _ip, _dp, C.POINTER(C.c_int)]),
    "oc_set_air_mesh": (C.c_int, [_ctx, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, C.c_double, C.c_double]),
    "oc_scaffold_update": (C.c_int, [_ctx, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_int)]),
    "oc_air_mesh_get": (C.c_int, [_ctx, _vp, _vp]),
    "oc_setup_times": (C.c_int, [_ctx, _dp]),
    "oc_amg_levels": (C.c_int, [_ctx, C.c_int, C.POINTER(C.c_int), _vp, _vp, _vp]),
    "oc_solver_set_pattern": (C.c_int, [_ctx, C.c_int, _ip, _ip, C.c_int, _ip, C.POINTER(C.c_int)]),
    "oc_solver_get_pattern": (C.c_int, [_ctx, _ip, _ip]),
    "oc_solver_update_a": (C.c_int, [_ctx, C.c_longlong, _ip, _ip, _dp]),
    "oc_solver_get_a": (C.c_int, [_ctx, _dp]),
    "oc_solver_set_a": (C.c_int, [_ctx, _dp]),
    "oc_solver_analyze_pattern": (C.c_int, [_ctx]),
    "oc_solver_factorize": (C.c_int, [_ctx]),
    "oc_solver_solve": (C.c_int, [_ctx, _dp, _dp]),
    "oc_solver_multiply": (C.c_int, [_ctx, _dp, _dp]),
    "oc_solver_set_params": (C.c_int, [_ctx, C.c_int, C.c_double, C.c_int]),
    "oc_solver_apply_precond": (C.c_int, [_ctx, _dp, _dp]),
    "oc_ic0_set_sweeps": (C.c_int, [_ctx, C.c_int, C.c_int]),
    "oc_amg_set_tail": (C.c_int, [_ctx, C.c_int]),
    "oc_solver_last_stats": (C.c_int, [_ctx, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "oc_solver_set_coords": (C.c_int, [_ctx, _dp]),
    "oc_solver_last_fallback": (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    "oc_newton_get_system": (C.c_int, [_ctx, _vp, _vp, _vp, _vp, _vp]),
    "oc_tutte": (C.c_int, [_ctx, C.c_int, C.c_int, C.c_double, C.c_int, _vp, C.POINTER(C.c_int),
                           C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "oc_set_charts": (C.c_int, [_ctx, C.c_int, _ip, _ip]),
    "oc_charts_state": (C.c_int, [_ctx, _vp, _vp, _vp, _vp]),
    "oc_newton_set_learned": (C.c_int, [_ctx, _vp]),
    "oc_gnn_num_params": (C.c_int, []),
    "oc_gnn_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "oc_gnn_destroy": (C.c_int, [_vp]),
    "oc_gnn_last_error": (C.c_char_p, [_vp]),
    "oc_gnn_set_params": (C.c_int, [_vp, _fp, C.c_int]),
    "oc_gnn_get_params": (C.c_int, [_vp, _fp, C.c_int]),
    "oc_gnn_get_grads": (C.c_int, [_vp, _fp, C.c_int]),
    "oc_gnn_set_system": (C.c_int, [_vp, C.c_int, _ip, _ip, _dp, _u8p]),
    "oc_gnn_forward": (C.c_int, [_vp, _vp]),
    "oc_gnn_zero_grad": (C.c_int, [_vp]),
    "oc_gnn_loss_grad": (C.c_int, [_vp, C.c_int, C.c_uint, C.POINTER(C.c_double)]),
    "oc_gnn_adam_step": (C.c_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]),
    "oc_gnn_precond_apply": (C.c_int, [_vp, _dp, _dp]),
    "oc_newton_set_precision": (C.c_int, [_ctx, C.c_int]),
    "oc_newton_set_refinement": (C.c_int, [_ctx, C.c_int]),
    "oc_newton_set_forcing": (C.c_int, [_ctx, C.c_double, C.c_double]),
    "oc_newton_init": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.POINTER(C.c_double)]),
    "oc_newton_step": (C.c_int, [_ctx, C.POINTER(IterStats), C.POINTER(C.c_int)]),
    "oc_newton_rebuild": (C.c_int, [_ctx, C.POINTER(C.c_double)]),
    "oc_time_kernel": (C.c_int, [_ctx, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "oc_system_sizes": (C.c_int, [_ctx] + [C.POINTER(C.c_longlong)] * 4),
    "oc_io_last_error": (C.c_char_p, []),
    "oc_obj_read_sizes": (C.c_int, [C.c_char_p] + [C.POINTER(C.c_int)] * 4),
    "oc_obj_read": (C.c_int, [C.c_char_p, _vp, _vp, _vp, _vp]),
    "oc_obj_write": (C.c_int, [C.c_char_p, C.c_int, _dp, C.c_int, _ip, C.c_int, _vp, _vp]),
    "oc_obj_write_mesh": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, C.c_int, _ip, _ip, C.c_int]),
    "oc_trace_open": (C.c_int, [_ctx, C.c_char_p]),
    "oc_trace_set_seam": (C.c_int, [_ctx, C.c_double]),
    "oc_trace_set_seams": (C.c_int, [_ctx, C.c_int, _ip, C.c_double, C.c_int]),
    "oc_info_write": (C.c_int, [_ctx, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                                _ip]),
    "oc_trace_close": (C.c_int, [_ctx]),
}

_lib = None


def header_functions():
    """Names of every function the public header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(oc_\w+)\s*\(", text, flags=re.M)))


def load():
    """Load the library (no GPU call is made here)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liboptcuts_gpu.so is not built (run optcuts_amd.build): " + LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _io_chk(st):
    if st != OC_OK:
        raise OCError(st, load().oc_io_last_error().decode())


def obj_read(path):
    """Read an OBJ mesh as igl::readOBJ does (main.cpp:1153). Returns (V (nV,3), UV (nUV,2), F (nF,3),
    FUV (nF,3) or None), indices 0-based. Host only: no GPU is used."""
    L = load()
    p = os.fsencode(path)
    nV, nUV, nF, has = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _io_chk(L.oc_obj_read_sizes(p, C.byref(nV), C.byref(nUV), C.byref(nF), C.byref(has)))
    V = np.zeros((nV.value, 3))
    UV = np.zeros((nUV.value, 2))
    F = np.zeros((nF.value, 3), np.int32)
    FUV = np.zeros((nF.value, 3), np.int32) if has.value else None
    _io_chk(L.oc_obj_read(p, V.ctypes.data, UV.ctypes.data, F.ctypes.data,
                          FUV.ctypes.data if FUV is not None else None))
    return V, UV, F, FUV


def obj_write(path, V, F, UV=None, FUV=None):
    """Write an OBJ mesh as TriMesh::save does (TriMesh.cpp:1758-1798)."""
    V, F = _f64(V), _i32(F)
    UV = _f64(UV) if UV is not None else np.zeros((0, 2))
    FUV = _i32(FUV) if FUV is not None else None
    _io_chk(load().oc_obj_write(os.fsencode(path), V.shape[0], V, F.shape[0], F, UV.shape[0], UV.ctypes.data,
                                FUV.ctypes.data if FUV is not None else None))


def obj_write_mesh(path, Vrest, UV, F, F0, scale_uv=False):
    """Write the cut mesh on the input mesh's vertices as TriMesh::saveAsMesh(path, F0, scaleUV) does
    (TriMesh.cpp:1939-1969)."""
    Vrest, UV, F, F0 = _f64(Vrest), _f64(UV), _i32(F), _i32(F0)
    _io_chk(load().oc_obj_write_mesh(os.fsencode(path), Vrest.shape[0], Vrest, UV, F.shape[0], F, F0,
                                     int(bool(scale_uv))))


class Context:
    """One device context (one parameterization problem) on GPU `device`."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        st = self.L.oc_ctx_create(device, C.byref(h))
        if st != OC_OK:
            raise OCError(st, "no usable GPU device %d" % device)
        self.h = h
        self.nV = self.nF = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.oc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st, ok=(OC_OK,)):
        if st not in ok:
            raise OCError(st, self.L.oc_last_error(self.h).decode())
        return st

    def arch(self):
        buf = C.create_string_buffer(64)
        self._chk(self.L.oc_device_arch(self.h, buf, 64))
        return buf.value.decode()

    # ---- mesh / Energy ------------------------------------------------------
    def set_mesh(self, Vrest, F, fixed=(0,), area_thres=0.0):
        Vrest, F = _f64(Vrest), _i32(F)
        fx = _i32(np.array(fixed, dtype=np.int32).reshape(-1))
        self._chk(self.L.oc_set_mesh(self.h, Vrest.shape[0], Vrest, F.shape[0], F, len(fx), fx, float(area_thres)))
        self.nV, self.nF = Vrest.shape[0], F.shape[0]

    def set_uv(self, UV):
        UV = _f64(UV)
        assert UV.shape == (self.nV, 2)
        self._chk(self.L.oc_set_uv(self.h, UV))

    def get_uv(self):
        out = np.zeros((self.nV, 2))
        self._chk(self.L.oc_get_uv(self.h, out))
        return out

    def mesh_stats(self):
        a, e = C.c_double(), C.c_double()
        self._chk(self.L.oc_mesh_stats(self.h, C.byref(a), C.byref(e)))
        return a.value, e.value

    def energy(self, uniform=False):
        e = C.c_double()
        self._chk(self.L.oc_energy(self.h, int(uniform), C.byref(e)))
        return e.value

    def energy_noalloc(self, uniform=False):
        """oc_energy without device allocation (the library's fixed scratch array)."""
        e = C.c_double()
        self._chk(self.L.oc_energy_noalloc(self.h, int(uniform), C.byref(e)))
        return e.value

    def energy_per_elem(self, uniform=False):
        out = np.zeros(self.nF)
        self._chk(self.L.oc_energy_per_elem(self.h, int(uniform), out))
        return out

    def gradient(self, uniform=False):
        out = np.zeros(2 * self.nV)
        self._chk(self.L.oc_gradient(self.h, int(uniform), out))
        return out

    def hessian_triplets(self, uniform=False):
        n = C.c_longlong()
        self._chk(self.L.oc_hessian_triplets(self.h, int(uniform), 0, None, None, None, C.byref(n)))
        I = np.zeros(n.value, np.int32)
        J = np.zeros(n.value, np.int32)
        V = np.zeros(n.value)
        self._chk(self.L.oc_hessian_triplets(self.h, int(uniform), n.value, I.ctypes.data, J.ctypes.data,
                                             V.ctypes.data, C.byref(n)))
        return I, J, V

    def element_terms(self, precision=PRECISION_FP64):
        """Per-triangle (E (nF,), G (nF, 6), H (nF, 6, 6)) of the element pass in the given precision."""
        E = np.zeros(self.nF)
        G = np.zeros((self.nF, 6))
        H = np.zeros((self.nF, 6, 6))
        self._chk(self.L.oc_element_terms(self.h, int(precision), E.ctypes.data, G.ctypes.data, H.ctypes.data))
        return E, G, H

    def init_step_size(self, d, step=1.0):
        s = C.c_double(step)
        self._chk(self.L.oc_init_step_size(self.h, _f64(d), C.byref(s)))
        return s.value

    def check_inversion(self):
        ok = C.c_int()
        self._chk(self.L.oc_check_inversion(self.h, C.byref(ok)))
        return bool(ok.value)

    def seam_sparsity(self, cohE, init_seam_len=0.0, tri_soup=False):
        """TriMesh::computeSeamSparsity at the current UV (cohE: (nCoh, 4) vertex ids, -1 on boundary edges)."""
        cohE = _i32(np.asarray(cohE).reshape(-1, 4))
        out = C.c_double()
        self._chk(self.L.oc_seam_sparsity(self.h, cohE.shape[0], cohE, float(init_seam_len), int(bool(tri_soup)),
                                          C.byref(out)))
        return out.value

    def pcg_graphs(self, enable):
        """PCG iterations from captured HIP graphs (True, the default) or launched one by one."""
        self._chk(self.L.oc_pcg_graphs(self.h, int(bool(enable))))

    def div_grad_per_vert(self):
        """SymDirichletEnergy::computeDivGradPerVert at the current UV: (nV,)."""
        out = np.zeros(self.nV)
        self._chk(self.L.oc_div_grad_per_vert(self.h, out))
        return out

    def split_candidates(self, split_interior=False, filter_exp_in=0.6, vert_weight=None):
        """TriMesh::querySplit's candidate filter (propagate = false): (candidates in key order, map size)."""
        w = None if vert_weight is None else _f64(vert_weight)
        cand = np.zeros(max(self.nV, 2), np.int32)
        n, ns = C.c_int(), C.c_int()
        self._chk(self.L.oc_split_candidates(self.h, int(split_interior), float(filter_exp_in),
                                             None if w is None else w.ctypes.data, cand.size, cand, C.byref(n),
                                             C.byref(ns)))
        return cand[:n.value].copy(), ns.value

    def merge_candidates(self, cohE):
        """TriMesh::queryMerge's per-pair screen: (status per row: 0 candidate, 1 not a zipper bottom,
        2 infeasible; merged positions (nCoh, 2); number of candidates)."""
        cohE = _i32(cohE).reshape(-1, 4)
        st = np.zeros(cohE.shape[0], np.int32)
        mp = np.zeros((cohE.shape[0], 2))
        n = C.c_int()
        self._chk(self.L.oc_merge_candidates(self.h, cohE.shape[0], cohE, st, mp.reshape(-1), C.byref(n)))
        return st, mp, n.value

    def set_air_mesh(self, airUV, Fair, bnd, area_thres, w_scaf):
        airUV, Fair, bnd = _f64(airUV), _i32(Fair), _i32(bnd)
        self._chk(self.L.oc_set_air_mesh(self.h, airUV.shape[0], airUV, Fair.shape[0], Fair, len(bnd), bnd,
                                         float(area_thres), float(w_scaf)))
        self.nVair, self.nFair = airUV.shape[0], Fair.shape[0]

    def scaffold_update(self, mode=SCAFFOLD_REMESH_WHEN_NEEDED, min_angle_deg=0.0, max_interval=8):
        """The air mesh between steps (rest shape reset to the current UV; Lawson re-mesh per mode: when needed =
        the last step's bound came from an air triangle, max_interval calls have passed since the last re-mesh
        (0: no count), or an air angle is below min_angle_deg (0: no angle test)). Returns (energy, edges
        flipped)."""
        e, n = C.c_double(), C.c_int()
        self._chk(self.L.oc_scaffold_update(self.h, int(mode), float(min_angle_deg), int(max_interval), C.byref(e),
                                            C.byref(n)))
        return e.value, n.value

    def amg_levels(self):
        """[(unknowns, blocks, block size)] per multigrid level of the last Newton system ([] without multigrid)."""
        cap = 32
        n = C.c_int()
        rows = np.zeros(cap, np.int64)
        blocks = np.zeros(cap, np.int64)
        bs = np.zeros(cap, np.int32)
        self._chk(self.L.oc_amg_levels(self.h, cap, C.byref(n), rows.ctypes.data, blocks.ctypes.data,
                                       bs.ctypes.data))
        return [(int(rows[l]), int(blocks[l]), int(bs[l])) for l in range(min(n.value, cap))]

    def setup_times(self):
        """Host ms of the last set-up phases (vertex lists, pattern+maps, upload, multigrid structure or its
        remap, learned graph, whole set-up) and of the last air-mesh update (download+flips, whole update), and
        amg_kept_remeshes: air re-meshes the current multigrid structure has been kept across."""
        ms = np.zeros(9)
        self._chk(self.L.oc_setup_times(self.h, ms))
        keys = ("vertex_lists", "pattern_maps", "upload", "amg_structure", "learned_graph", "setup_total",
                "air_flips", "air_update_total", "amg_kept_remeshes")
        return dict(zip(keys, ms.tolist()))

    def air_mesh_get(self):
        """(airUV (nVair, 2), Fair (nFair, 3)) of the current air mesh."""
        uv = np.zeros((self.nVair, 2))
        F = np.zeros((self.nFair, 3), np.int32)
        self._chk(self.L.oc_air_mesh_get(self.h, uv.ctypes.data, F.ctypes.data))
        return uv, F

    # ---- LinSysSolver --------------------------------------------------------
    def solver_set_pattern(self, nV, nbr_ptr, nbr, fixed=(0,)):
        fx = _i32(np.array(fixed, dtype=np.int32).reshape(-1))
        nnz = C.c_int()
        self._chk(self.L.oc_solver_set_pattern(self.h, nV, _i32(nbr_ptr), _i32(nbr), len(fx), fx, C.byref(nnz)))
        self._sol_n, self._sol_nnz = 2 * nV, nnz.value
        ia = np.zeros(2 * nV + 1, np.int32)
        ja = np.zeros(nnz.value, np.int32)
        self._chk(self.L.oc_solver_get_pattern(self.h, ia, ja))
        return ia, ja

    def solver_update_a(self, I, J, V):
        self._chk(self.L.oc_solver_update_a(self.h, len(I), _i32(I), _i32(J), _f64(V)))

    def solver_set_a(self, a):
        a = _f64(a)
        assert a.shape == (self._sol_nnz,)
        self._chk(self.L.oc_solver_set_a(self.h, a))

    def solver_get_a(self):
        a = np.zeros(self._sol_nnz)
        self._chk(self.L.oc_solver_get_a(self.h, a))
        return a

    def solver_set_params(self, precond=PRECOND_JACOBI, rel_tol=1e-10, max_iter=100000):
        self._chk(self.L.oc_solver_set_params(self.h, int(precond), float(rel_tol), int(max_iter)))

    def solver_analyze_pattern(self):
        self._chk(self.L.oc_solver_analyze_pattern(self.h))

    def solver_factorize(self):
        self._chk(self.L.oc_solver_factorize(self.h))

    def solver_solve(self, rhs, allow_not_converged=False):
        x = np.zeros(self._sol_n)
        ok = (OC_OK, -7) if allow_not_converged else (OC_OK,)
        self._chk(self.L.oc_solver_solve(self.h, _f64(rhs), x), ok)
        return x

    def solver_multiply(self, x):
        y = np.zeros(self._sol_n)
        self._chk(self.L.oc_solver_multiply(self.h, _f64(x), y))
        return y

    def solver_last_stats(self):
        it, rr = C.c_int(), C.c_double()
        self._chk(self.L.oc_solver_last_stats(self.h, C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def solver_set_coords(self, UV):
        """Coordinates (nV, 2) for the multigrid's rigid near-null space (PRECOND_AMG)."""
        UV = _f64(UV)
        assert UV.size == self._sol_n
        self._chk(self.L.oc_solver_set_coords(self.h, UV))

    def solver_apply_precond(self, r):
        """z = M^-1 r with the drop-in's factorized preconditioner (PRECOND_AMG or PRECOND_IC0)."""
        r = _f64(r)
        assert r.size == self._sol_n
        z = np.zeros(self._sol_n)
        self._chk(self.L.oc_solver_apply_precond(self.h, r, z))
        return z

    def amg_set_tail(self, max_nodes=-1):
        """Largest last smoothed multigrid level run as one single-workgroup launch (0: off, -1: built-in limit);
        takes effect at the next multigrid structure set-up."""
        self._chk(self.L.oc_amg_set_tail(self.h, int(max_nodes)))

    def ic0_set_sweeps(self, factor_sweeps=3, apply_sweeps=2):
        """PRECOND_IC0's fixed-point factor sweeps and Jacobi sweeps per triangular solve."""
        self._chk(self.L.oc_ic0_set_sweeps(self.h, int(factor_sweeps), int(apply_sweeps)))

    def solver_last_fallback(self):
        """0 none, 1 multigrid rebuilt with the Gershgorin interval, 2 Jacobi-PCG."""
        f = C.c_int()
        self._chk(self.L.oc_solver_last_fallback(self.h, C.byref(f)))
        return f.value

    # ---- Newton driver -------------------------------------------------------
    def tutte(self, layout=TUTTE_COMPONENTS, precond=PRECOND_JACOBI, rel_tol=1e-12, max_iter=200000):
        """Tutte initial map of the current mesh (becomes the context's UV). Returns (UV, loops pinned,
        vertices pinned, PCG iterations)."""
        out = np.zeros((self.nV, 2))
        nl, nb, it = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.L.oc_tutte(self.h, int(layout), int(precond), float(rel_tol), int(max_iter),
                                  out.ctypes.data, C.byref(nl), C.byref(nb), C.byref(it)))
        return out, nl.value, nb.value, it.value

    def trace_open(self, folder):
        """Append per-iteration lines to folder/energyValPerIter.txt and folder/gradientPerIter.txt."""
        self._chk(self.L.oc_trace_open(self.h, os.fsencode(folder)))

    def trace_set_seam(self, seam):
        self._chk(self.L.oc_trace_set_seam(self.h, float(seam)))

    def trace_close(self):
        self._chk(self.L.oc_trace_close(self.h))

    def trace_set_seams(self, cohE, init_seam_len=0.0, fracture_mode=False):
        """Seam column recomputed per energy line from the cohesive edges (Optimizer.cpp:706-728)."""
        cohE = _i32(np.asarray(cohE).reshape(-1, 4))
        self._chk(self.L.oc_trace_set_seams(self.h, cohE.shape[0], cohE, float(init_seam_len),
                                            int(bool(fracture_mode))))

    def info_write(self, path, vert_amt_input, nF_initial, iter_num, topo_iter, lambda_init, timings,
                   init_seams=None):
        """info.txt as saveInfoForPresent writes it (main.cpp:346-387); timings: 15 values."""
        t = _f64(np.asarray(timings, dtype=np.float64).reshape(-1))
        assert t.size == 15
        s = _i32(np.zeros((0, 4)) if init_seams is None else np.asarray(init_seams).reshape(-1, 4))
        self._chk(self.L.oc_info_write(self.h, os.fsencode(path), int(vert_amt_input), int(nF_initial),
                                       int(iter_num), int(topo_iter), float(lambda_init), t, s.shape[0],
                                       s if s.size else np.zeros(4, np.int32)))

    def set_charts(self, vertex_start, face_start):
        """Batched charts: contiguous vertex / triangle ranges of the mesh set by set_mesh."""
        vs, fs = _i32(vertex_start), _i32(face_start)
        self._chk(self.L.oc_set_charts(self.h, len(vs) - 1, vs, fs))
        self.n_charts = len(vs) - 1

    def charts_state(self):
        n = self.n_charts
        E, g, a = np.zeros(n), np.zeros(n), np.zeros(n)
        st = np.zeros(n, np.int32)
        self._chk(self.L.oc_charts_state(self.h, E.ctypes.data, g.ctypes.data, a.ctypes.data, st.ctypes.data))
        return {"E_SD": E, "grad_sq": g, "step": a, "state": st}

    def set_learned(self, net):
        """Bind a Gnn for PRECOND_LEARNED (None unbinds); the context keeps a reference."""
        self._learned = net
        self._chk(self.L.oc_newton_set_learned(self.h, net.h if net is not None else None))

    def set_forcing(self, loose_tol=0.0, step_below=0.0):
        """After a Newton step shorter than step_below (and for the first step), solve to max(pcg_rel_tol,
        loose_tol); loose_tol 0 turns it off."""
        self._chk(self.L.oc_newton_set_forcing(self.h, float(loose_tol), float(step_below)))

    def set_refinement(self, max_passes):
        """Mixed-precision iterative refinement of the Newton solve (PCG on the FP32-rounded matrix, FP64
        residual), at most max_passes passes; 0 turns it off."""
        self._chk(self.L.oc_newton_set_refinement(self.h, int(max_passes)))

    def newton_init(self, a0=1.0, rel_gl2_tol=1e-12, allow_edec_rel_tol=True, precond=PRECOND_JACOBI,
                    pcg_rel_tol=1e-10, pcg_max_iter=20000, precision=PRECISION_FP64):
        self._chk(self.L.oc_newton_set_precision(self.h, int(precision)))
        e = C.c_double()
        self._chk(self.L.oc_newton_init(self.h, float(a0), float(rel_gl2_tol), int(allow_edec_rel_tol),
                                        int(precond), float(pcg_rel_tol), int(pcg_max_iter), C.byref(e)))
        return e.value

    def newton_rebuild(self):
        """After set_air_mesh: the scaffold rebuild of Optimizer::solve (new system, same parameters, no trace
        line). Returns the energy at the current UV."""
        e = C.c_double()
        self._chk(self.L.oc_newton_rebuild(self.h, C.byref(e)))
        return e.value

    def newton_step(self):
        s = IterStats()
        done = C.c_int()
        self._chk(self.L.oc_newton_step(self.h, C.byref(s), C.byref(done)))
        return s.as_dict(), bool(done.value)

    def time_kernel(self, which, reps=20):
        ms = C.c_double()
        self._chk(self.L.oc_time_kernel(self.h, int(which), int(reps), C.byref(ms)))
        return ms.value

    def get_system(self):
        """The last Newton system: (rowPtr, colIdx, vals (nnzb x 2 x 2), rhs (2 nV), fixed mask)."""
        _, nV, nnzb, _ = self.system_sizes()
        rp = np.zeros(nV + 1, np.int32)
        ci = np.zeros(max(nnzb, 1), np.int32)
        va = np.zeros(4 * max(nnzb, 1))
        b = np.zeros(2 * nV)
        fm = np.zeros(nV, np.uint8)
        self._chk(self.L.oc_newton_get_system(self.h, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, b.ctypes.data,
                                              fm.ctypes.data))
        return rp, ci[:nnzb], va[:4 * nnzb].reshape(-1, 2, 2), b, fm

    def system_sizes(self):
        v = [C.c_longlong() for _ in range(4)]
        self._chk(self.L.oc_system_sizes(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)


class Gnn:
    """Learned preconditioner network on GPU `device` (see include/optcuts_gpu.h)."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        st = self.L.oc_gnn_create(device, C.byref(h))
        if st != OC_OK:
            raise OCError(st, "no usable GPU device %d" % device)
        self.h = h
        self.n_params = self.L.oc_gnn_num_params()
        self.nV = 0
        self.rp = None

    def close(self):
        if getattr(self, "h", None):
            self.L.oc_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != OC_OK:
            raise OCError(st, self.L.oc_gnn_last_error(self.h).decode())

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._chk(self.L.oc_gnn_set_params(self.h, t, t.size))

    def get_params(self):
        t = np.zeros(self.n_params, np.float32)
        self._chk(self.L.oc_gnn_get_params(self.h, t, t.size))
        return t

    def get_grads(self):
        t = np.zeros(self.n_params, np.float32)
        self._chk(self.L.oc_gnn_get_grads(self.h, t, t.size))
        return t

    def set_system(self, rp, ci, vals, fixed):
        rp, ci = _i32(rp), _i32(ci)
        self._chk(self.L.oc_gnn_set_system(self.h, len(rp) - 1, rp, ci, _f64(np.asarray(vals).reshape(-1)),
                                           np.ascontiguousarray(fixed, dtype=np.uint8)))
        self.nV = len(rp) - 1
        self.rp, self.ci = rp, ci

    def forward(self):
        """Factor per CSR block (nnzb x 2 x 2, float32)."""
        out = np.zeros(4 * len(self.ci), np.float32)
        self._chk(self.L.oc_gnn_forward(self.h, out.ctypes.data))
        return out.reshape(-1, 2, 2)

    def zero_grad(self):
        self._chk(self.L.oc_gnn_zero_grad(self.h))

    def loss_grad(self, n_probes, seed):
        loss = C.c_double()
        self._chk(self.L.oc_gnn_loss_grad(self.h, int(n_probes), int(seed) & 0xFFFFFFFF, C.byref(loss)))
        return loss.value

    def adam_step(self, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
        self._chk(self.L.oc_gnn_adam_step(self.h, float(lr), float(beta1), float(beta2), float(eps), int(t)))

    def precond_apply(self, r):
        r = _f64(r)
        z = np.zeros_like(r)
        self._chk(self.L.oc_gnn_precond_apply(self.h, r, z))
        return z
