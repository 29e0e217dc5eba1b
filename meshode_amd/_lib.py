"""Loader for libmeshode.so (the HIP/gfx950 C ABI declared in include/meshode.h).

There is no CPU fallback: if the library is missing or no GPU is visible,
calls raise.  Build the library with ``__graft_entry__.build()`` or
``make -C meshode_amd/csrc``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmeshode.so")

# Exported symbols (kept in sync with include/meshode.h; tests check both).
SYMBOLS = [
    "mo_version", "mo_last_error",
    "mo_template_create_f32", "mo_template_create_f64", "mo_template_destroy", "mo_template_info",
    "mo_template_grid", "mo_grid_build",
    "mo_normalize_by_template", "mo_denormalize_by_template",
    "mo_distance_loss_forward", "mo_distance_loss_backward",
    "mo_rigidity_store", "mo_rigid_loss_forward", "mo_rigid_loss_backward",
    "mo_graph_store", "mo_graph_loss_forward", "mo_graph_loss_backward",
    "mo_lm_default_options", "mo_deform_batch", "mo_rigid_deform_batch", "mo_last_timings",
    "mo_lm_shape_count", "mo_lm_shape_info", "mo_lm_set_shape_table",
    "mo_cad_prepare", "mo_cad_sizes", "mo_cad_export", "mo_cad_destroy", "mo_delaunay2", "mo_delaunay3",
    "mo_apply_transform_f64", "mo_linear_estimation", "mo_deform_subdivision",
    "mo_linear_estimation_rot", "mo_nearest_f32", "mo_deform_rot", "mo_deform_rot_batch",
    "mo_ode_rk4_workspace_bytes", "mo_ode_rk4_forward", "mo_ode_rk4_backward",
    "mo_flow256_packed_floats", "mo_flow256_pack", "mo_flow256_step", "mo_flow256_step_save",
    "mo_flow256_step_backward", "mo_flow256_small_grad_floats", "mo_flow256_small_grads",
    "mo_flow256_dw_work_floats", "mo_flow256_dw_grads",
    "mo_deform_batch_noalloc_max_bytes", "mo_deform_batch_noalloc", "mo_template_rebuild_f64",
    "mo_last_grid_stats", "mo_rigid_deform_batch_graph", "mo_batch_graph_launch", "mo_batch_graph_destroy",
    "mo_rigid_deform_batch_async", "mo_batch_wait",
]


class MoError(RuntimeError):
    """A libmeshode call returned a negative status."""

    def __init__(self, func, code, msg):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


class LMOptions(C.Structure):
    _fields_ = [
        ("max_num_iterations", C.c_int32),
        ("max_consecutive_invalid_steps", C.c_int32),
        ("function_tolerance", C.c_double),
        ("gradient_tolerance", C.c_double),
        ("parameter_tolerance", C.c_double),
        ("min_relative_decrease", C.c_double),
        ("initial_trust_region_radius", C.c_double),
        ("max_trust_region_radius", C.c_double),
        ("min_trust_region_radius", C.c_double),
        ("min_lm_diagonal", C.c_double),
        ("max_lm_diagonal", C.c_double),
        ("pcg_max_iterations", C.c_int32),
        ("pcg_relative_tolerance", C.c_float),
        ("block_threads", C.c_int32),
        ("timing", C.c_int32),
        ("kernel_variant", C.c_int32),
        ("max_successful_steps", C.c_int32),
        ("team_vertices", C.c_int32),
    ]


class LMStats(C.Structure):
    _fields_ = [
        ("initial_cost", C.c_double),
        ("final_cost", C.c_double),
        ("vertex_cost", C.c_double),
        ("rigidity_cost", C.c_double),
        ("iterations", C.c_int32),
        ("successful_steps", C.c_int32),
        ("unsuccessful_steps", C.c_int32),
        ("termination", C.c_int32),
        ("pcg_iterations", C.c_int64),
        ("clock_start", C.c_int64),
        ("clock_end", C.c_int64),
        ("solver_path", C.c_int32),
        ("workgroups", C.c_int32),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["solver_path_name"] = SOLVER_PATHS.get(self.solver_path, "unknown")
        return d


# mo_lm_stats.solver_path (MO_PATH_* in include/meshode.h)
SOLVER_PATHS = {0: "lds", 1: "one_workgroup", 2: "team", 3: "one_workgroup_coop_refused", 4: "lds_team"}


_lib = None
_lock = threading.Lock()

_vp, _i32, _i64, _f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
_SIGS = {
    "mo_version": (C.c_int, []),
    "mo_last_error": (C.c_char_p, []),
    "mo_template_create_f32": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _i32, _vp, C.POINTER(_i32)]),
    "mo_template_create_f64": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _i32, _vp, C.POINTER(_i32)]),
    "mo_template_destroy": (C.c_int, [_i32]),
    "mo_template_rebuild_f64": (C.c_int, [_i32, _vp, _i64, _vp, _i64, _vp]),
    "mo_template_info": (C.c_int, [_i32, C.POINTER(_i32), C.POINTER(_f64), _vp]),
    "mo_template_grid": (C.c_int, [_i32, _vp, _vp, _vp, _vp]),
    "mo_grid_build": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _vp, _vp, _vp]),
    "mo_normalize_by_template": (C.c_int, [_vp, _i64, _i32, _vp]),
    "mo_denormalize_by_template": (C.c_int, [_vp, _i64, _i32, _vp]),
    "mo_distance_loss_forward": (C.c_int, [_vp, _i64, _i32, _vp, _vp]),
    "mo_distance_loss_backward": (C.c_int, [_vp, _i64, _i32, _vp, _vp]),
    "mo_rigidity_store": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp]),
    "mo_rigid_loss_forward": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _vp]),
    "mo_rigid_loss_backward": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _vp]),
    "mo_graph_store": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp]),
    "mo_graph_loss_forward": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _vp]),
    "mo_graph_loss_backward": (C.c_int, [_vp, _i64, _vp, _i64, _i32, _vp, _vp]),
    "mo_lm_default_options": (None, [C.POINTER(LMOptions)]),
    "mo_deform_batch": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(LMOptions), _vp, _vp]),
    "mo_deform_batch_noalloc_max_bytes": (_i64, []),
    "mo_deform_batch_noalloc": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(LMOptions), _vp,
                                          _vp]),
    "mo_rigid_deform_batch": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                        C.POINTER(LMOptions), _vp, _vp, _vp]),
    "mo_rigid_deform_batch_graph": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                              C.POINTER(LMOptions), _vp, _vp, _vp, C.POINTER(_vp)]),
    "mo_batch_graph_launch": (C.c_int, [_vp, _vp, _vp]),
    "mo_batch_graph_destroy": (C.c_int, [_vp]),
    "mo_rigid_deform_batch_async": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                              C.POINTER(LMOptions), _vp, _vp, _vp, C.POINTER(_vp)]),
    "mo_batch_wait": (C.c_int, [_vp, _vp]),
    "mo_last_timings": (C.c_int, [_vp]),
    "mo_last_grid_stats": (C.c_int, [_vp]),
    "mo_lm_shape_count": (_i32, []),
    "mo_lm_shape_info": (C.c_int, [_i32, C.POINTER(_i32), C.POINTER(_i32)]),
    "mo_lm_set_shape_table": (C.c_int, [_vp, _vp, _i32]),
    "mo_cad_prepare": (C.c_int, [_vp, _i64, _vp, _i64, _f64, _f64, _f64, _i32, C.POINTER(_i64)]),
    "mo_cad_sizes": (C.c_int, [_i64, _vp]),
    "mo_cad_export": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mo_cad_destroy": (C.c_int, [_i64]),
    "mo_delaunay2": (C.c_int, [_vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "mo_delaunay3": (C.c_int, [_vp, _i64, _vp, _i64, C.POINTER(_i64), C.POINTER(_i32)]),
    "mo_apply_transform_f64": (C.c_int, [_vp, _i64, _i32, _i32, _vp]),
    "mo_linear_estimation": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _f64, C.POINTER(_i32), _vp]),
    "mo_linear_estimation_rot": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _f64, C.POINTER(_i32), _vp]),
    "mo_nearest_f32": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _vp, _vp]),
    "mo_ode_rk4_workspace_bytes": (_i64, [_i64]),
    "mo_flow256_packed_floats": (_i64, []),
    "mo_flow256_pack": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mo_flow256_step": (C.c_int, [_vp, _i64, _vp, C.c_float, _vp, _vp]),
    "mo_flow256_step_save": (C.c_int, [_vp, _i64, _vp, C.c_float, _vp, _vp, _vp, _vp]),
    "mo_flow256_step_backward": (C.c_int, [_i64, _vp, C.c_float, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mo_flow256_small_grad_floats": (_i64, []),
    "mo_flow256_small_grads": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mo_flow256_dw_work_floats": (_i64, []),
    "mo_flow256_dw_grads": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "mo_ode_rk4_forward": (C.c_int, [_vp, _i64, _vp, _vp, C.c_float, C.c_float, _vp, _vp, _vp]),
    "mo_ode_rk4_backward": (C.c_int, [_vp, _i64, _vp, _vp, C.c_float, C.c_float, _vp, _vp, _vp, _vp, _vp, _i64,
                                      _vp]),
    "mo_deform_rot": (C.c_int, [_i32, _vp, _i64, _vp, _i64, _f64, C.POINTER(LMOptions), _vp,
                                C.POINTER(LMStats), _vp]),
    "mo_deform_rot_batch": (C.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(LMOptions), _vp,
                                      _vp, _vp]),
    "mo_deform_subdivision": (C.c_int, [_i32, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _f64, C.POINTER(LMOptions), _i32,
                                        _vp, _vp, _vp, C.POINTER(_i32), _vp]),
}

# void (*)(const double* V, int64_t nv, void* user)
FlowCallback = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_void_p)


def load(path: str = LIB_PATH):
    """Load libmeshode.so (no GPU needed to load; calls need one)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise FileNotFoundError(
                    f"{path} not found: build it with `make -C meshode_amd/csrc` "
                    "(or __graft_entry__.build()); there is no CPU fallback")
            lib = C.CDLL(path)
            for name, (res, args) in _SIGS.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def call(name, *args):
    """Call an int-returning entry point and raise MoError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise MoError(name, rc, lib.mo_last_error().decode(errors="replace"))
    return rc


def last_timings():
    """Phase times of this thread's last batch call made with options.timing=1:
    dict(grid_ms, setup_ms, lm_ms, total_ms, clock_start, clock_hz)."""
    import numpy as np
    out = np.zeros(6)
    call("mo_last_timings", C.c_void_p(out.ctypes.data))
    return dict(grid_ms=out[0], setup_ms=out[1], lm_ms=out[2], total_ms=out[3], clock_start=out[4],
                clock_hz=out[5])


def last_grid_stats():
    """Grid-build work counters of this thread's last rigid_deform batch call
    made with options.timing=2: dict(exact_tests, exact_tests_needed,
    bound_checks, tiles) (lane-level counts, summed over the batch)."""
    import numpy as np
    out = np.zeros(4, np.int64)
    call("mo_last_grid_stats", C.c_void_p(out.ctypes.data))
    return dict(exact_tests=int(out[0]), exact_tests_needed=int(out[1]), bound_checks=int(out[2]),
                tiles=int(out[3]))


def default_lm_options() -> LMOptions:
    o = LMOptions()
    load().mo_lm_default_options(C.byref(o))
    return o
