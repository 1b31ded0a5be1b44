"""ctypes binding of ``libtfnr.so`` (the C ABI declared in include/tf_nr.h).

The library is built in-tree by ``tissuefolding_paperversion_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtfnr.so")

TF_OK, TF_NOT_CONVERGED, TF_KRYLOV_STALL, TF_NEG_JACOBIAN, TF_NAN, TF_OOM, TF_HIP_ERROR, TF_INVALID = range(8)
STATUS_NAMES = {0: "TF_OK", 1: "TF_NOT_CONVERGED", 2: "TF_KRYLOV_STALL", 3: "TF_NEG_JACOBIAN", 4: "TF_NAN",
                5: "TF_OOM", 6: "TF_HIP_ERROR", 7: "TF_INVALID"}
KRYLOV_AUTO, KRYLOV_BICGSTAB, KRYLOV_CG, KRYLOV_CG1, KRYLOV_PIPECG = -1, 0, 1, 3, 4
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
PRECOND_AUTO, PRECOND_BLOCK_JACOBI, PRECOND_TWO_LEVEL = -1, 0, 1

# Every symbol include/tf_nr.h declares (checked by the CPU test suite).
EXPORTS = ["tf_create", "tf_destroy", "tf_last_error", "tf_default_options", "tf_set_options", "tf_get_options",
           "tf_get_sizes", "tf_set_positions", "tf_get_positions", "tf_set_growth_rates", "tf_set_fg", "tf_get_fg",
           "tf_set_material", "tf_grow", "tf_update_viscosity_surfaces", "tf_get_viscosity_surfaces",
           "tf_begin_step", "tf_assemble", "tf_get_residual", "tf_get_jacobian_bsr", "tf_element_dump", "tf_solve",
           "tf_update_uk", "tf_nr_step", "tf_timestep", "tf_get_segment_stats", "tf_kernel_timing",
           "tf_kernel_bytes", "tf_kernel_probe", "tf_set_tiling", "tf_get_tiling", "tf_autotune", "tf_set_bound_dofs",
           "tf_set_ablated_nodes", "tf_set_growth_state", "tf_get_growth_rotation", "tf_apply_jacobian",
           "tf_solve_system", "tf_adjoint_elements", "tf_adjoint_reset", "tf_adjoint_step", "tf_adjoint_get",
           "tf_get_node_masses", "tf_detect_packing", "tf_set_packing_pairs", "tf_get_packing_pairs",
           "tf_set_growth_increments", "tf_element_dump_range_noalloc", "tf_get_coarse_operator",
           "tf_kernel_probe_named", "tf_krylov_timing", "tf_copy_bandwidth", "tf_set_growth_functions",
           "tf_pin_growth_positions", "tf_set_time", "tf_get_time", "tf_get_growth_increments"]


class TfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_elements", C.c_int32), ("node_pos", C.c_void_p), ("conn", C.c_void_p),
                ("ref_pos", C.c_void_p), ("lam", C.c_void_p), ("mu", C.c_void_p), ("eta_int", C.c_void_p),
                ("eta_ext", C.c_void_p), ("fixed", C.c_void_p), ("exposed", C.c_void_p), ("growth_rate", C.c_void_p),
                ("n_segments", C.c_int32), ("node_seg_ptr", C.c_void_p), ("elem_seg_ptr", C.c_void_p),
                ("node_placement", C.c_void_p)]


class SolverOpts(C.Structure):
    _fields_ = [("krylov", C.c_int32), ("krylov_max_iter", C.c_int32), ("krylov_rtol", C.c_double),
                ("krylov_atol", C.c_double), ("iters_per_sync", C.c_int32),
                ("nr_tol", C.c_double), ("nr_max_iter", C.c_int32), ("tangent_precision", C.c_int32),
                ("matrix_precision", C.c_int32), ("growth_rotation", C.c_int32), ("use_graph", C.c_int32),
                ("krylov_forcing", C.c_int32), ("xcd_tile_map", C.c_int32), ("krylov_warm_start", C.c_int32),
                ("implicit_packing", C.c_int32), ("preconditioner", C.c_int32), ("coarse_aggregates", C.c_int32),
                ("coarse_refresh", C.c_int32), ("column_blocks", C.c_int32), ("precond_fp64", C.c_int32),
                ("mixed_fp64_from", C.c_int32), ("nr_graph", C.c_int32), ("tl_persistent", C.c_int32),
                ("tl_barrier_polls", C.c_int32)]


class GrowthFunction(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("t0", C.c_double), ("t1", C.c_double),
                ("rate", C.c_double * 3), ("angle", C.c_double), ("centre", C.c_double * 2), ("inner", C.c_double),
                ("outer", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
                ("grid_rates", C.POINTER(C.c_double)), ("grid_angles", C.POINTER(C.c_double))]


class StepStats(C.Structure):
    _fields_ = [("nr_iterations", C.c_int32), ("krylov_iterations", C.c_int32), ("converged", C.c_int32),
                ("rotated_elements", C.c_int32), ("krylov_fallbacks", C.c_int32), ("packing_pairs", C.c_int32),
                ("last_du_norm", C.c_double), ("ms_grow", C.c_double),
                ("ms_assemble", C.c_double), ("ms_solve", C.c_double), ("ms_total", C.c_double),
                ("krylov_by_newton", C.c_int32 * 4), ("two_level", C.c_int32), ("krylov_stalls", C.c_int32),
                ("host_waits", C.c_int32), ("nr_graph_builds", C.c_int32), ("krylov_persistent", C.c_int32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["krylov_by_newton"] = list(d["krylov_by_newton"])
        return d


_lib = None


def load(path: str | None = None):
    """Load libtfnr.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m tissuefolding_paperversion_amd.build` "
                           "(there is no CPU fallback)")
    L = C.CDLL(p)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "tf_create": ([C.POINTER(MeshDesc), C.POINTER(vp)], i32),
        "tf_destroy": ([vp], None),
        "tf_last_error": ([vp], C.c_char_p),
        "tf_default_options": ([C.POINTER(SolverOpts)], None),
        "tf_set_options": ([vp, C.POINTER(SolverOpts)], i32),
        "tf_get_options": ([vp, C.POINTER(SolverOpts)], i32),
        "tf_get_sizes": ([vp, vp, vp, vp, vp], i32),
        "tf_set_positions": ([vp, vp], i32),
        "tf_get_positions": ([vp, vp], i32),
        "tf_set_growth_rates": ([vp, vp], i32),
        "tf_set_fg": ([vp, vp], i32),
        "tf_get_fg": ([vp, vp, vp], i32),
        "tf_set_material": ([vp, vp, vp, vp], i32),
        "tf_grow": ([vp, dbl, vp], i32),
        "tf_update_viscosity_surfaces": ([vp], i32),
        "tf_get_viscosity_surfaces": ([vp, vp], i32),
        "tf_begin_step": ([vp], i32),
        "tf_assemble": ([vp, dbl], i32),
        "tf_get_residual": ([vp, vp], i32),
        "tf_get_jacobian_bsr": ([vp, vp, vp, vp], i32),
        "tf_element_dump": ([vp, i32, dbl, vp, vp, vp, vp], i32),
        "tf_element_dump_range_noalloc": ([vp, i32, i32, dbl, vp, vp, vp, vp], i32),
        "tf_get_coarse_operator": ([vp, vp, vp, vp], i32),
        "tf_solve": ([vp, vp, vp, vp], i32),
        "tf_update_uk": ([vp, vp, vp], i32),
        "tf_nr_step": ([vp, dbl, C.POINTER(StepStats)], i32),
        "tf_timestep": ([vp, dbl, C.POINTER(StepStats)], i32),
        "tf_get_segment_stats": ([vp, vp, vp, vp], i32),
        "tf_kernel_timing": ([vp, C.c_char_p, i32, vp, vp], i32),
        "tf_kernel_bytes": ([vp, C.c_char_p, vp], i32),
        "tf_kernel_probe": ([vp, i32, vp], i32),
        "tf_kernel_probe_named": ([vp, C.c_char_p, i32, vp], i32),
        "tf_krylov_timing": ([vp, i32, vp, vp, vp], i32),
        "tf_copy_bandwidth": ([vp, C.c_int64, i32, vp], i32),
        "tf_set_growth_functions": ([vp, i32, vp, vp, vp, i32, i32], i32),
        "tf_pin_growth_positions": ([vp], i32),
        "tf_set_time": ([vp, dbl], i32),
        "tf_get_time": ([vp, vp], i32),
        "tf_get_growth_increments": ([vp, vp], i32),
        "tf_set_tiling": ([vp, i32, i32], i32),
        "tf_get_tiling": ([vp, vp, vp, vp], i32),
        "tf_autotune": ([vp, dbl, i32, vp, vp, vp, vp], i32),
        "tf_set_bound_dofs": ([vp, i32, vp], i32),
        "tf_set_ablated_nodes": ([vp, i32, vp], i32),
        "tf_set_growth_state": ([vp, vp, vp], i32),
        "tf_get_growth_rotation": ([vp, vp], i32),
        "tf_apply_jacobian": ([vp, i32, vp, vp], i32),
        "tf_solve_system": ([vp, i32, vp, vp, vp, vp], i32),
        "tf_adjoint_elements": ([vp, vp, vp, vp, vp], i32),
        "tf_adjoint_reset": ([vp], i32),
        "tf_adjoint_step": ([vp, dbl, vp, vp, vp, vp, vp], i32),
        "tf_adjoint_get": ([vp, vp, vp, vp, vp], i32),
        "tf_get_node_masses": ([vp, vp], i32),
        "tf_detect_packing": ([vp, vp], i32),
        "tf_set_packing_pairs": ([vp, i32, vp, vp, vp], i32),
        "tf_get_packing_pairs": ([vp, vp, vp, vp, vp, vp], i32),
        "tf_set_growth_increments": ([vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)
