"""ctypes binding of the C oracle (``oracle/build/liblpbf_oracle.so``).

TEST INFRASTRUCTURE ONLY: this is the checker that the tests, the smoke check
and the benchmark's CPU-baseline leg compare the HIP path against.  The C
functions restate the reference (file:line cited in ``lpbf_oracle.c``).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liblpbf_oracle.so")
REF_PROBE_PATH = os.path.join(HERE, "_ref", "libref_headers.so")


class StepConst(C.Structure):
    """Reference include/lpbf_step_const.hpp:21-32 (216 bytes)."""
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
    _fields_ = [("x_l", C.c_double), ("y_l", C.c_double),
                ("q_scale", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("tmax", C.c_double), ("arg", C.c_longlong),
                ("ymin", C.c_int), ("ymax", C.c_int),
                ("molten", C.c_longlong), ("sum_dT", C.c_double),
                ("sum_src", C.c_double), ("sum_loss", C.c_double)]


class OraMaterial(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k", "rho", "cp", "Tm", "eps", "eta")]


class OraProcess(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("P", "V", "r0", "h", "hConv", "dSub", "Tamb", "Tsub")] + [
        ("radiation", C.c_bool)]


class OraDomain(C.Structure):
    _fields_ = [("Lx", C.c_double), ("Ly", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int)]


class OraTick(C.Structure):
    _fields_ = [("step", C.c_int), ("t", C.c_double), ("x_l", C.c_double),
                ("y_l", C.c_double), ("q_scale", C.c_double),
                ("x_laser_csv", C.c_double), ("wrapped", C.c_int),
                ("dump", C.c_int)]


class OraMetrics(C.Structure):
    _fields_ = [("peak_T", C.c_double), ("W_um", C.c_double),
                ("W_sub_um", C.c_double), ("x_hot", C.c_int)]


class OraRaster(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double),
                ("length", C.c_double), ("hatch", C.c_double),
                ("n_tracks", C.c_int), ("V", C.c_double)]


_DP = C.POINTER(C.c_double)


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    d, i, v = C.c_double, C.c_int, None
    sig = {
        "ora_defaults": (v, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "ora_apply_case": (i, [C.c_char_p, C.c_void_p]),
        "ora_dt_cfl": (d, [C.c_void_p, C.c_void_p]),
        "ora_dt_cfl_ext": (d, [C.c_void_p, C.c_void_p, d]),
        "ora_k_of": (d, [d]), "ora_cp_of": (d, [d]),
        "ora_k_enh": (d, [d, d]), "ora_k_eff": (d, [d, d]),
        "ora_cp_eff": (d, [d]), "ora_alpha_eff_max": (d, [d, d]),
        "ora_make_const": (v, [C.c_void_p, C.c_void_p, C.c_void_p, d, C.c_void_p]),
        "ora_step": (v, [_DP, _DP, C.c_void_p, i]),
        "ora_step_rows": (v, [_DP, _DP, C.c_void_p, i, i, i]),
        "ora_step_rows_f32": (v, [C.c_void_p, C.c_void_p, C.c_void_p, i, i]),
        "ora_step_ex": (v, [_DP, _DP, C.c_void_p, i, i, d, d, C.c_void_p]),
        "ora_step_tile": (v, [_DP, _DP, i, i, C.c_void_p, i]),
        "ora_step_tiled": (v, [_DP, _DP, C.c_void_p, i, i]),
        "ora_tile_bounds": (v, [i, i, C.c_void_p, C.c_void_p]),
        "ora_make_tile": (v, [_DP, i, i, i, i, _DP]),
        "ora_measure": (v, [_DP, i, i, d, d, C.c_void_p]),
        "ora_wsub_from_col": (d, [_DP, i, d, d]),
        "ora_reduce_partials": (v, [_DP, i, i, d, i, C.c_void_p]),
        "ora_reduce_finalize": (v, [C.c_void_p, i, i, i, C.c_void_p]),
        "ora_build_plan": (v, [d, d, d, d, i, i, C.c_void_p]),
        "ora_track_plan": (v, [d, d, d, d, i, C.c_void_p]),
        "ora_raster_plan": (v, [C.c_void_p, d, i, C.c_void_p]),
        "ora_noise": (d, [C.c_uint64, C.c_uint64]),
        "ora_noise_field": (v, [C.c_uint64, d, d, _DP, C.c_longlong]),
        "ora_run": (v, [C.c_void_p, C.c_void_p, C.c_void_p, d, i, i, d, i,
                        _DP, _DP, _DP, _DP, _DP]),
        "ora_run_plan": (v, [C.c_void_p, i, i, d, d, C.c_void_p, i, _DP,
                             C.c_void_p]),
        "ora_step_tile_f32": (v, [C.c_void_p, C.c_void_p, i, i, C.c_void_p]),
        "ora_steptask_run": (i, [i, C.c_void_p, i, C.c_void_p, i, i, i, C.c_void_p,
                                 C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


@dataclass
class Setup:
    mat: OraMaterial
    proc: OraProcess
    dom: OraDomain


def defaults(case: str | None = "0") -> Setup:
    m, p, d = OraMaterial(), OraProcess(), OraDomain()
    lib().ora_defaults(C.byref(m), C.byref(p), C.byref(d))
    if case is not None and lib().ora_apply_case(case.encode(), C.byref(p)) != 0:
        raise ValueError(f"unknown case '{case}'")
    return Setup(m, p, d)


def dt_cfl(s: Setup) -> float:
    return lib().ora_dt_cfl(C.byref(s.mat), C.byref(s.dom))


def dt_cfl_ext(s: Setup, kmult: float = 1.0) -> float:
    return lib().ora_dt_cfl_ext(C.byref(s.mat), C.byref(s.dom), kmult)


def make_const(s: Setup, dt: float) -> StepConst:
    k = StepConst()
    lib().ora_make_const(C.byref(s.mat), C.byref(s.proc), C.byref(s.dom), dt,
                         C.byref(k))
    return k


def step(T: np.ndarray, k: StepConst, phys: int = 0) -> np.ndarray:
    out = np.empty_like(T)
    lib().ora_step(_dp(T), _dp(out), C.byref(k), phys)
    return out


def step_threads_f32(T: np.ndarray, k: StepConst, threads: int = 1,
                     out: np.ndarray | None = None, pool=None) -> np.ndarray:
    """ora_step_rows_f32 (baseline physics, mirror edges) on a float32 field,
    rows split over `threads` host threads."""
    assert T.dtype == np.float32 and T.flags.c_contiguous
    out = np.empty_like(T) if out is None else out
    ny = T.shape[0]
    cuts = [ny * t // threads for t in range(threads + 1)]
    pT, pO, pk = T.ctypes.data, out.ctypes.data, C.byref(k)

    def part(t):
        lib().ora_step_rows_f32(pT, pO, pk, cuts[t], cuts[t + 1])

    if pool is None or threads == 1:
        for t in range(threads):
            part(t)
    else:
        list(pool.map(part, range(threads)))
    return out


def step_threads(T: np.ndarray, k: StepConst, phys: int = 0, threads: int = 1,
                 out: np.ndarray | None = None, pool=None) -> np.ndarray:
    """ora_step with the rows split over `threads` host threads (the ctypes
    calls release the GIL); equals step() exactly."""
    out = np.empty_like(T) if out is None else out
    ny = T.shape[0]
    cuts = [ny * t // threads for t in range(threads + 1)]
    pT, pO, pk = _dp(T), _dp(out), C.byref(k)

    def part(t):
        lib().ora_step_rows(pT, pO, pk, phys, cuts[t], cuts[t + 1])

    if pool is None or threads == 1:
        for t in range(threads):
            part(t)
    else:
        list(pool.map(part, range(threads)))
    return out


def steptask_run(T: np.ndarray, k: StepConst, plan: np.ndarray, n_steps: int, ntiles: int,
                 nthreads: int, phys: int = 0):
    """ora_steptask_run: n_steps of the reference CPU solver's loop (tiles
    built, stepped by a thread pool, stitched back) on T in place; float64 or
    float32 (baseline physics).  Returns the seconds spent building tiles,
    waiting for and stitching results, and in total."""
    assert T.flags.c_contiguous and T.dtype in (np.float64, np.float32)
    p = np.ascontiguousarray(plan, dtype=np.float64)
    assert p.shape[0] >= n_steps and p.shape[1] == 3
    ph = (C.c_double * 3)()
    rc = lib().ora_steptask_run(32 if T.dtype == np.float32 else 64, C.byref(k), phys,
                                p.ctypes.data, n_steps, ntiles, nthreads, T.ctypes.data, ph)
    if rc != 0:
        raise RuntimeError("ora_steptask_run failed")
    return tuple(ph)


def step_ex(T: np.ndarray, k: StepConst, phys: int = 0, bc: int = 0,
            T_bc: float = 0.0, Tm: float = 1609.0):
    out = np.empty_like(T)
    st = StepStats()
    lib().ora_step_ex(_dp(T), _dp(out), C.byref(k), phys, bc, T_bc, Tm,
                      C.byref(st))
    return out, st


def step_tiled(T: np.ndarray, k: StepConst, phys: int, ntiles: int) -> np.ndarray:
    out = np.empty_like(T)
    lib().ora_step_tiled(_dp(T), _dp(out), C.byref(k), phys, ntiles)
    return out


def step_tile(tile: np.ndarray, owned: int, y0: int, k: StepConst,
              phys: int = 0) -> np.ndarray:
    out = np.empty_like(tile)
    lib().ora_step_tile(_dp(tile), _dp(out), owned, y0, C.byref(k), phys)
    return out


def make_tile(T: np.ndarray, y0: int, y1: int) -> np.ndarray:
    ny, nx = T.shape
    tile = np.empty((y1 - y0 + 3, nx))
    lib().ora_make_tile(_dp(T), nx, ny, y0, y1, _dp(tile))
    return tile


def tile_bounds(ny: int, ntiles: int):
    y0 = (C.c_int * ntiles)()
    y1 = (C.c_int * ntiles)()
    lib().ora_tile_bounds(ny, ntiles, y0, y1)
    return list(y0), list(y1)


def measure(T: np.ndarray, dy: float, Tm: float) -> OraMetrics:
    ny, nx = T.shape
    m = OraMetrics()
    lib().ora_measure(_dp(np.ascontiguousarray(T)), nx, ny, dy, Tm, C.byref(m))
    return m


def wsub_from_col(col: np.ndarray, dy: float, Tm: float) -> float:
    return lib().ora_wsub_from_col(_dp(np.ascontiguousarray(col)), col.size, dy, Tm)


def reduce_partials(T: np.ndarray, Tm: float, nblocks: int):
    ny, nx = T.shape
    parts = (Partial * nblocks)()
    lib().ora_reduce_partials(_dp(np.ascontiguousarray(T)), nx, ny, Tm, nblocks, parts)
    return parts


def reduce_finalize(parts, nx: int, ny: int) -> Partial:
    out = Partial()
    lib().ora_reduce_finalize(parts, len(parts), nx, ny, C.byref(out))
    return out


def build_plan(V: float, dt: float, Lx: float, Ly: float, n_steps: int,
               snap_every: int):
    plan = (OraTick * n_steps)()
    lib().ora_build_plan(V, dt, Lx, Ly, n_steps, snap_every, plan)
    return plan


def track_plan(V, dt, Lx, Ly, n_steps) -> np.ndarray:
    plan = np.zeros((n_steps, 3))
    lib().ora_track_plan(V, dt, Lx, Ly, n_steps, plan.ctypes.data_as(C.c_void_p))
    return plan


def raster_plan(r: OraRaster, dt: float, n_steps: int) -> np.ndarray:
    plan = np.zeros((n_steps, 3))
    lib().ora_raster_plan(C.byref(r), dt, n_steps, plan.ctypes.data_as(C.c_void_p))
    return plan


def noise(seed: int, idx: int) -> float:
    return lib().ora_noise(seed, idx)


def noise_field(seed: int, base: float, amp: float, shape) -> np.ndarray:
    T = np.empty(shape)
    lib().ora_noise_field(seed, base, amp, _dp(T), T.size)
    return T


def run(s: Setup, dt: float, n_steps: int, phys: int = 0, kmult: float = 1.0,
        ntiles: int = 0, T0: np.ndarray | None = None):
    nx, ny = s.dom.nx, s.dom.ny
    T = np.full((ny, nx), s.proc.Tamb) if T0 is None else np.array(T0, dtype=np.float64)
    pk, W, Ws = (np.zeros(n_steps) for _ in range(3))
    summ = np.zeros(3)
    lib().ora_run(C.byref(s.mat), C.byref(s.proc), C.byref(s.dom), dt, n_steps,
                  phys, kmult, ntiles, _dp(T), _dp(pk), _dp(W), _dp(Ws), _dp(summ))
    return T, pk, W, Ws, summ


def run_plan(k: StepConst, plan: np.ndarray, T0: np.ndarray, phys: int = 0,
             bc: int = 0, T_bc: float = 0.0, Tm: float = 1609.0):
    T = np.array(T0, dtype=np.float64, order="C")
    n = plan.shape[0]
    stats = (StepStats * n)()
    p = np.ascontiguousarray(plan, dtype=np.float64)
    lib().ora_run_plan(C.byref(k), phys, bc, T_bc, Tm,
                       p.ctypes.data_as(C.c_void_p), n, _dp(T), stats)
    return T, stats


def ref_probe() -> C.CDLL | None:
    """The reference's own header arithmetic, if oracle/_ref was built."""
    if not os.path.exists(REF_PROBE_PATH):
        return None
    lib_ = C.CDLL(REF_PROBE_PATH)
    d, i = C.c_double, C.c_int
    lib_.ref_make_const.argtypes = [_DP, _DP, i, d, d, i, i, d, C.c_void_p]
    lib_.ref_make_const.restype = None
    lib_.ref_dt_cfl.argtypes = [_DP, d, d, i, i]
    lib_.ref_dt_cfl.restype = d
    lib_.ref_dt_cfl_ext.argtypes = [_DP, d, d, i, i, d]
    lib_.ref_dt_cfl_ext.restype = d
    for n in ("ref_k_of", "ref_cp_of", "ref_cp_eff"):
        getattr(lib_, n).argtypes = [d]
        getattr(lib_, n).restype = d
    for n in ("ref_k_enh", "ref_k_eff", "ref_alpha_eff_max"):
        getattr(lib_, n).argtypes = [d, d]
        getattr(lib_, n).restype = d
    lib_.ref_sizeof_stepconst.restype = i
    lib_.ref_sizeof_partial.restype = i
    return lib_
