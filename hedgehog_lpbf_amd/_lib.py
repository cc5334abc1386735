"""ctypes binding of libhedgehog_lpbf_amd.so (the gfx950 step library).

The library is the product: every compute call below runs the HIP kernels.
Loading fails loudly if the shared library has not been built; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhedgehog_lpbf_amd.so")

LPBF_BC_MIRROR = 0
LPBF_BC_DIRICHLET = 1
LPBF_PHYS_BASELINE = 0
LPBF_PHYS_EXTENDED = 1
LPBF_ERR_INVALID_ARGUMENT = 1000
LPBF_ERR_SCRATCH_LIMIT = 1001
LPBF_F32_NOALLOC_SCRATCH_BYTES = 4 << 20


class StepConst(C.Structure):
    """include/lpbf_step_const.h (reference include/lpbf_step_const.hpp:21-32)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("inv_dx2", C.c_double),
        ("inv_dy2", C.c_double), ("inv_r0_sq", C.c_double),
        ("alpha", C.c_double), ("K", C.c_double), ("rho", C.c_double),
        ("cp", C.c_double), ("rho_cp", C.c_double), ("inv_h", C.c_double),
        ("h", C.c_double), ("dt", C.c_double),
        ("peak_S", C.c_double), ("eta", C.c_double), ("P", C.c_double),
        ("hConv", C.c_double), ("Tamb", C.c_double), ("Tsub", C.c_double),
        ("dSub", C.c_double), ("eps", C.c_double),
        ("radiation", C.c_bool),
        ("kmult", C.c_double),
        ("x_l", C.c_double), ("y_l", C.c_double),
        ("q_scale", C.c_double),
    ]


class Partial(C.Structure):
    _fields_ = [("tmax", C.c_double), ("arg", C.c_longlong),
                ("ymin", C.c_int), ("ymax", C.c_int)]


class LaserTick(C.Structure):
    _fields_ = [("x_l", C.c_double), ("y_l", C.c_double), ("q_scale", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("tmax", C.c_double), ("arg", C.c_longlong),
                ("ymin", C.c_int), ("ymax", C.c_int),
                ("molten", C.c_longlong), ("sum_dT", C.c_double),
                ("sum_src", C.c_double), ("sum_loss", C.c_double)]


class RunOptions(C.Structure):
    """include/lpbf_run.h lpbf_run_options."""
    _fields_ = [("precision", C.c_int), ("bc", C.c_int), ("T_bc", C.c_double),
                ("batch", C.c_int), ("depth", C.c_int), ("rows", C.c_int),
                ("T_ref", C.c_double), ("keep_records", C.c_int), ("graph", C.c_int),
                ("f64_depth", C.c_int), ("cruise_cols", C.c_int), ("pad_", C.c_int)]


class StepRecord(C.Structure):
    """include/lpbf_run.h lpbf_step_record."""
    _fields_ = [("stats", StepStats), ("W_um", C.c_double), ("W_sub_um", C.c_double),
                ("step", C.c_int), ("has_col", C.c_int)]


class F32Instance(C.Structure):
    """include/lpbf_batch.h lpbf_f32_instance (device pointers as integers)."""
    _fields_ = [("k", StepConst), ("theta0", C.c_void_p), ("theta1", C.c_void_p),
                ("ld", C.c_int), ("size_class", C.c_int), ("T_ref", C.c_double),
                ("T_bc", C.c_double), ("Tm", C.c_double), ("d_plan", C.c_void_p),
                ("d_stats", C.c_void_p), ("n_steps", C.c_int), ("pad_", C.c_int)]


class F32ColJob(C.Structure):
    """include/lpbf_batch.h lpbf_f32_col_job (device pointers as integers)."""
    _fields_ = [("theta", C.c_void_p * 2), ("stats", C.c_void_p), ("cruise", C.c_void_p),
                ("col", C.c_void_p), ("T_ref", C.c_double), ("ld", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("n_steps", C.c_int)]


assert C.sizeof(F32ColJob) == 64
assert C.sizeof(StepConst) == 216 and C.sizeof(Partial) == 24
assert C.sizeof(StepRecord) == 80 and C.sizeof(F32Instance) == 288
assert C.sizeof(LaserTick) == 24 and C.sizeof(StepStats) == 56


class MrWindow(C.Structure):
    """include/lpbf_multirate.h lpbf_mr_window."""
    _fields_ = [("fine", C.c_void_p), ("c0", C.c_void_p), ("c1", C.c_void_p),
                ("plan", C.c_void_p), ("dx", C.c_double), ("dy", C.c_double),
                ("inv_r0_sq", C.c_double), ("ldf", C.c_int), ("fx", C.c_int), ("fy", C.c_int),
                ("ox", C.c_int), ("oy", C.c_int), ("ldc", C.c_int), ("nxc", C.c_int),
                ("nyc", C.c_int), ("step", C.c_int), ("ld1", C.c_int), ("c1x", C.c_int),
                ("c1y", C.c_int), ("pad_", C.c_int)]


class MrConfig(C.Structure):
    """include/lpbf_multirate.h lpbf_mr_config."""
    _fields_ = [("k", StepConst), ("T_ref", C.c_double), ("tol_K", C.c_double),
                ("n_steps", C.c_int), ("recheck_every", C.c_int), ("overlap", C.c_int),
                ("use_network", C.c_int), ("coarse_depth", C.c_int), ("coarse_rows", C.c_int)]


class MrCounters(C.Structure):
    """include/lpbf_multirate.h lpbf_mr_counters."""
    _fields_ = [("blocks", C.c_longlong), ("calls", C.c_longlong), ("applied", C.c_longlong),
                ("rechecks", C.c_longlong), ("fallbacks", C.c_longlong),
                ("fd_blocks", C.c_longlong), ("max_recheck_K", C.c_double),
                ("max_applied_recheck_K", C.c_double)]


assert C.sizeof(MrWindow) == 112 and C.sizeof(MrConfig) == 256 and C.sizeof(MrCounters) == 64

_vp, _i, _d, _v = C.c_void_p, C.c_int, C.c_double, None
_ll, _ull, _sz = C.c_longlong, C.c_ulonglong, C.c_size_t
_DP = C.POINTER(C.c_double)
_KP = C.POINTER(StepConst)   # host StepConst: pass the struct itself, which ctypes
                             # keeps alive for the duration of the call

# name -> (restype, argtypes); pointers to device memory and streams are
# passed as integers (c_void_p).
SIGNATURES = {
    # reference seam (include/lpbf_kernels.h)
    "lpbf_launch_step": (_v, [_vp, _vp, _KP, _i, _vp]),
    "lpbf_launch_reduce": (_v, [_vp, _i, _i, _d, _vp, _i, _vp]),
    "lpbf_launch_step_reduce": (_v, [_vp, _vp, _KP, _i, _d, _vp, _vp]),
    "lpbf_launch_reduce_finalize": (_v, [_vp, _i, _i, _i, _vp, _vp]),
    "lpbf_launch_extract_col": (_v, [_vp, _i, _i, _vp, _vp, _vp]),
    # status
    "lpbf_get_last_error": (_i, []),
    "lpbf_error_string": (C.c_char_p, [_i]),
    # FP64 additions
    "lpbf_launch_step_tile": (_v, [_vp, _vp, _i, _i, _KP, _i, _vp]),
    "lpbf_launch_step_ex": (_v, [_vp, _vp, _KP, _i, _i, _d, _d, _vp, _vp]),
    "lpbf_step_ex_nblocks": (_i, [_i, _i]),
    "lpbf_launch_stats_finalize": (_v, [_vp, _i, _i, _i, _vp, _vp]),
    "lpbf_step_ex_plan_nblocks": (_i, [_i, _i]),
    "lpbf_launch_step_ex_plan": (_v, [_vp, _vp, _KP, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _i, _vp, _vp]),
    "lpbf_launch_noise_f64": (_v, [_vp, _ll, _ull, _d, _d, _vp]),
    "lpbf_probe_exp_f64": (_v, [_vp, _vp, _ll, _vp]),
    "lpbf_step_ex_batch_nblocks": (_i, [_i, _i]),
    "lpbf_launch_step_ex_batch": (_v, [_vp, _vp, _KP, _i, _i, _d, _d, _vp, _i, _vp, _vp, _vp, _vp,
                                       _vp]),
    "lpbf_launch_step_ex_batch_flush": (_v, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "lpbf_step_ensemble_nblocks": (_i, [_i, _i, _i]),
    "lpbf_step_ensemble_rows": (_i, [_i, _i, _i]),
    "lpbf_launch_step_ensemble": (_v, [_vp, _i, _i, _i, _i, _i, _i, _d, _d, _vp, _i, _vp]),
    "lpbf_launch_step_ensemble_flush": (_v, [_vp, _i, _i, _i, _i, _vp, _i, _vp]),
    "lpbf_f64_blocked_workspace_bytes": (C.c_size_t, [_i, _i, _i, _i]),
    "lpbf_launch_step_f64_blocked": (_v, [_vp, _vp, _KP, _i, _i, _d, _d, _vp, _i, _i, _i, _vp,
                                          _vp, _vp]),
    "lpbf_f64_blocked_autotune": (_i, [_KP, _i, _i, _d, _d, _vp, _i, _i, _vp, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    # FP32 path
    "lpbf_f32_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "lpbf_launch_step_f32": (_v, [_vp, _vp, _i, _KP, _d, _i, _d, _d, _vp, _i, _i, _i,
                                  _vp, _vp, _vp]),
    "lpbf_launch_step_f32_noalloc": (_v, [_vp, _vp, _i, _KP, _d, _i, _d, _d, _vp, _i, _i, _i,
                                          _vp, _vp]),
    "lpbf_launch_f32_stats_finalize": (_v, [_vp, _i, _i, _i, _i, _vp, _vp]),
    "lpbf_f32_autotune": (_i, [_KP, _d, _i, _d, _d, _vp, _i, _i, _vp, C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "lpbf_launch_noise_f32": (_v, [_vp, _i, _i, _i, _ull, _d, _d, _vp]),
    "lpbf_launch_theta_to_T": (_v, [_vp, _i, _i, _i, _d, _vp, _vp]),
    "lpbf_launch_T_to_theta": (_v, [_vp, _i, _i, _d, _vp, _i, _vp]),
    "lpbf_launch_extract_col_f32": (_v, [_vp, _i, _i, _i, _d, _vp, _vp, _vp]),
    # batched FP32 instances (include/lpbf_batch.h)
    "lpbf_f32_batch_create": (_i, [C.POINTER(F32Instance), _i, _i, C.POINTER(_vp)]),
    "lpbf_f32_batch_destroy": (_v, [_vp]),
    "lpbf_f32_batch_set_tiling": (_i, [_vp, _i, _i, _i]),
    "lpbf_f32_batch_get_tiling": (_i, [_vp, _i, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    "lpbf_f32_batch_autotune": (_i, [_vp, _i, _vp, C.POINTER(C.c_double)]),
    "lpbf_f32_batch_enqueue": (_i, [_vp, _i, _i, _vp]),
    "lpbf_f32_batch_capture": (_i, [_vp, _i, _i, _vp]),
    "lpbf_f32_batch_launch_graph": (_i, [_vp, _vp]),
    "lpbf_f32_batch_current": (_i, [_vp, _i]),
    "lpbf_f32_ext_workspace_bytes": (C.c_size_t, [_i, _i]),
    "lpbf_launch_step_f32_ext": (_v, [_vp, _vp, _i, _KP, _d, _i, _d, _d, _vp, _i, _vp, _vp, _vp]),
    "lpbf_f32_batch_enqueue_launch": (_i, [_vp, _i, _i, _vp]),
    "lpbf_launch_hot_cols_f32": (_i, [_vp, _i, _i, _i, _i, C.c_longlong, _vp]),
    "lpbf_f32_batch_launches": (_i, [_vp, _i, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    # surrogate (include/lpbf_surrogate.h)
    "lpbf_sg_pack_bytes": (_sz, []),
    "lpbf_sg_work_bytes": (_sz, [_i, _i]),
    "lpbf_sg_pack": (_v, [_vp, _vp, _vp]),
    "lpbf_sg_forward": (_v, [_vp, _vp, _i, _vp, _vp, _vp]),
    "lpbf_sg_loss_grad": (_v, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "lpbf_sg_adam": (_v, [_vp, _vp, _vp, _vp, _i, _i, C.c_float, C.c_float, C.c_float,
                          C.c_float, _vp]),
    # two-level run (include/lpbf_multirate.h)
    "lpbf_mr_restrict": (_v, [_vp, _i, _i, _i, _vp, _i, _vp]),
    "lpbf_mr_window_origins": (_i, [_vp, _i, _i, _i, _d, _d, _vp, _vp]),
    "lpbf_mr_coarse_coeffs": (_i, [_KP, _d, _vp]),
    "lpbf_mr_gather": (_v, [_vp, _i, _vp, _vp, _vp, _vp]),
    "lpbf_mr_target": (_v, [_vp, _i, _vp, _vp, _vp, _vp]),
    "lpbf_mr_create": (_i, [C.POINTER(MrConfig), _vp, _vp, _vp, C.POINTER(C.c_void_p)]),
    "lpbf_mr_destroy": (_v, [_vp]),
    "lpbf_mr_advance": (_i, [_vp, _i]),
    "lpbf_mr_field": (_i, [_vp, _vp]),
    "lpbf_mr_get_counters": (_i, [_vp, C.POINTER(MrCounters)]),
    "lpbf_mr_recheck_diffs": (_i, [_vp, _vp, _i]),
    # device-resident run (include/lpbf_run.h)
    "lpbf_run_create": (_i, [_KP, _i, _d, _d, _d, _d, _vp, _vp, _vp, _vp, _i,
                             C.POINTER(RunOptions), C.POINTER(_vp)]),
    "lpbf_run_destroy": (_v, [_vp]),
    "lpbf_run_upload": (_i, [_vp, _vp]),
    "lpbf_run_noise": (_i, [_vp, _ull, _d, _d]),
    "lpbf_run_advance": (_i, [_vp, _i]),
    "lpbf_run_download": (_i, [_vp, _vp]),
    "lpbf_run_summary": (_i, [_vp, _DP, C.POINTER(C.c_int)]),
    "lpbf_run_records": (_i, [_vp, _i, _i, C.POINTER(StepRecord)]),
    "lpbf_run_graph_launches": (_i, [_vp]),
    # ensemble of runs on one plate (include/lpbf_run.h)
    "lpbf_ensemble_create": (_i, [_KP, _i, _i, _d, _d, _d, _vp, _vp, _vp, _vp, _vp,
                                  C.POINTER(RunOptions), C.POINTER(_vp)]),
    "lpbf_ensemble_destroy": (_v, [_vp]),
    "lpbf_ensemble_upload": (_i, [_vp, _i, _vp]),
    "lpbf_ensemble_advance": (_i, [_vp, _i]),
    "lpbf_ensemble_download": (_i, [_vp, _i, _vp]),
    "lpbf_ensemble_summary": (_i, [_vp, _i, _DP, C.POINTER(C.c_int)]),
    "lpbf_ensemble_graph_launches": (_i, [_vp]),
    "lpbf_ensemble_records": (_i, [_vp, _i, _i, _i, C.POINTER(StepRecord)]),
    # host helpers (include/lpbf_host.h)
    "lpbf_make_const": (_v, [_DP, _DP, _i, _d, _d, _i, _i, _d, _KP]),
    "lpbf_dt_cfl": (_d, [_DP, _d, _d, _i, _i]),
    "lpbf_dt_cfl_ext": (_d, [_DP, _d, _d, _i, _i, _d]),
    "lpbf_plan_single_track": (_v, [_d, _d, _d, _d, _i, _vp, _vp, _vp]),
    "lpbf_plan_raster": (_v, [_d, _d, _d, _d, _i, _d, _d, _i, _vp]),
}

_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        # torch bundles its own HIP runtime (soname libamdhip64.so.7, loaded
        # through the name libamdhip64.so); loaded first, it is the runtime
        # this library binds to as well.  Loaded after this library, it would
        # come in as a second runtime next to the system one, and device
        # memory from one is unusable by the other.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
                "`make -C hedgehog_lpbf_amd/csrc`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _LIB = h
    return _LIB


class LpbfError(RuntimeError):
    pass


def check() -> None:
    """Raise if any lpbf_* call recorded an error since the last check."""
    err = lib().lpbf_get_last_error()
    if err:
        raise LpbfError(f"lpbf error {err}: {lib().lpbf_error_string(err).decode()}")
