"""ctypes binding of libpaintop_hip.so (the C ABI in include/paintop_hip.h).

This is the host side of the drop-in boundary: the structs mirror the header
field for field, and every call goes through the C ABI. There is no CPU
fallback: if the library is missing, loading raises ``HipLibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpaintop_hip.so")

PH_OK, PH_ERROR, PH_UNAVAILABLE = 0, 1, 2
PH_OP_POISSON_BLEND, PH_OP_SCREENED_POISSON = 0, 1
PH_FLAG_TIMING, PH_FLAG_WARM_START_NET, PH_FLAG_FORCE_PLAIN = 0x1, 0x2, 0x4
STOP_NAMES = {0: "converged", 1: "max-iterations", 2: "stalled"}
# paintop_ir::ErrorClass in declaration order (paintop-ir/src/error.rs:23-37)
ERROR_CLASSES = ("parse", "schema", "reference", "type", "semantic", "policy", "execution")

# Every symbol the header declares; the library must export all of them.
EXPORTED_SYMBOLS = (
    "ph_abi_version",
    "ph_probe",
    "ph_force_unavailable",
    "ph_poisson_solve",
    "ph_poisson_solve_device",
    "ph_poisson_solve_device_noalloc",
    "ph_fbm_device",
    "ph_polygon_mask_device",
    "ph_ellipse_mask_device",
    "ph_gauss_noise_device",
    "ph_poisson_solve_batch_device",
    "ph_unet_param_storage",
    "ph_unet_param_count",
    "ph_unet_create",
    "ph_unet_destroy",
    "ph_unet_forward",
    "ph_unet_train_step",
    "ph_unet_get_state",
    "ph_unet_set_state",
    "ph_warm_start_set_weights",
    "ph_warm_coarsen_device",
    "ph_poisson_solve_start_device",
    "ph_unet_gather_device",
    "ph_assert_no_change_outside_mask_device",
    "ph_coarsen_plane_device",
    "ph_correction_input_device",
    "ph_sweep_config",
)


MAX_LOCATIONS = 256  # PH_MAX_LOCATIONS (assert.rs:72)


class PhOutsideScan(ctypes.Structure):
    _fields_ = [
        ("changed_outside", ctypes.c_uint64),
        ("max_outside", ctypes.c_double),
        ("worst_x", ctypes.c_int64),
        ("worst_y", ctypes.c_int64),
        ("worst_delta", ctypes.c_double),
        ("n_locations", ctypes.c_uint32),
        ("reserved_", ctypes.c_uint32),
        ("locations", (ctypes.c_int64 * 2) * MAX_LOCATIONS),
    ]


class HipLibraryMissing(RuntimeError):
    """libpaintop_hip.so is not built; run ``python -m paintop_amd.build``."""


class PhSolveDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("channels", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("tolerance", ctypes.c_double),
        ("lam", ctypes.c_double),
        ("omega", ctypes.c_double),
        ("flags", ctypes.c_uint32),
        ("half_sweeps", ctypes.c_uint32),
        ("op", ctypes.c_int32),
        ("tile_rows", ctypes.c_uint32),
    ]


class PhReport(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("stop_reason", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("history_len", ctypes.c_uint32),
        ("tolerance", ctypes.c_double),
        ("final_residual", ctypes.c_double),
        ("residual_history", ctypes.POINTER(ctypes.c_double)),
        ("history_capacity", ctypes.c_uint32),
        ("worst_channel", ctypes.c_uint32),
        ("refinements", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
        ("unknowns", ctypes.c_uint64),
        ("unknown_sweeps", ctypes.c_uint64),
        ("device_ms", ctypes.c_double),
        ("omega_used", ctypes.c_double),
        ("half_sweeps_used", ctypes.c_uint32),
        ("warm_start_used", ctypes.c_uint32),
        ("warm_start_rel", ctypes.c_double),
        ("sweep_ms", ctypes.c_double),
        ("coarse_corrections", ctypes.c_uint32),
        ("reserved_", ctypes.c_uint32),
    ]


class PhBatchItem(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("channels", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("source", ctypes.c_void_p),
        ("target", ctypes.c_void_p),
        ("mask", ctypes.c_void_p),
        ("out_candidate", ctypes.c_void_p),
    ]


class PhError(ctypes.Structure):
    _fields_ = [
        ("error_class", ctypes.c_int32),
        ("code", ctypes.c_char * 64),
        ("message", ctypes.c_char * 512),
    ]


class PhDeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 128),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int32),
        ("device_index", ctypes.c_int32),
        ("runtime_version", ctypes.c_int32),
        ("driver_version", ctypes.c_int32),
        ("hbm_bytes", ctypes.c_uint64),
    ]


class PhTrainOpts(ctypes.Structure):
    _fields_ = [
        ("learning_rate", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("weight_decay", ctypes.c_float),
        ("apply_update", ctypes.c_int32),
        ("pyramid_weight", ctypes.c_float),
    ]


class HipCallError(RuntimeError):
    """A failed C-ABI call, carrying the paintop error class and code."""

    def __init__(self, status: int, error_class: str, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.status = status
        self.error_class = error_class
        self.code = code
        self.message = message


_lib = None


def lib():
    """Load libpaintop_hip.so (once). Raises HipLibraryMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipLibraryMissing(f"{LIB_PATH} is missing: run `python -m paintop_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.ph_abi_version.restype = ctypes.c_int32
    L.ph_probe.argtypes = [ctypes.POINTER(PhDeviceInfo), ctypes.POINTER(PhError)]
    L.ph_probe.restype = ctypes.c_int
    L.ph_force_unavailable.argtypes = [ctypes.c_int32]
    L.ph_poisson_solve.argtypes = [ctypes.POINTER(PhSolveDesc), vp, vp, vp, vp,
                                   ctypes.POINTER(PhReport), ctypes.POINTER(PhError)]
    L.ph_poisson_solve.restype = ctypes.c_int
    L.ph_poisson_solve_device.argtypes = [ctypes.POINTER(PhSolveDesc), vp, vp, vp, vp, vp,
                                          ctypes.POINTER(PhReport), ctypes.POINTER(PhError)]
    L.ph_poisson_solve_device.restype = ctypes.c_int
    L.ph_poisson_solve_device_noalloc.argtypes = [ctypes.POINTER(PhSolveDesc), vp, vp, vp, vp,
                                                  vp, ctypes.POINTER(PhReport),
                                                  ctypes.POINTER(PhError)]
    L.ph_poisson_solve_device_noalloc.restype = ctypes.c_int
    L.ph_poisson_solve_start_device.argtypes = [ctypes.POINTER(PhSolveDesc), vp, vp, vp, vp, vp,
                                                vp, ctypes.POINTER(PhReport),
                                                ctypes.POINTER(PhError)]
    L.ph_poisson_solve_start_device.restype = ctypes.c_int
    L.ph_fbm_device.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                ctypes.c_float, ctypes.c_float, vp, ctypes.POINTER(PhError)]
    L.ph_fbm_device.restype = ctypes.c_int
    L.ph_polygon_mask_device.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_uint32,
                                         ctypes.c_int32, vp, ctypes.POINTER(PhError)]
    L.ph_polygon_mask_device.restype = ctypes.c_int
    L.ph_ellipse_mask_device.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, vp,
                                         ctypes.POINTER(PhError)]
    L.ph_ellipse_mask_device.restype = ctypes.c_int
    L.ph_assert_no_change_outside_mask_device.argtypes = [
        vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
        ctypes.c_double, ctypes.c_double, vp, ctypes.POINTER(PhOutsideScan),
        ctypes.POINTER(PhError)]
    L.ph_assert_no_change_outside_mask_device.restype = ctypes.c_int
    L.ph_coarsen_plane_device.argtypes = ([vp, vp] + [ctypes.c_int32] * 9 +
                                          [vp, vp, ctypes.POINTER(PhError)])
    L.ph_coarsen_plane_device.restype = ctypes.c_int
    L.ph_correction_input_device.argtypes = ([vp, vp, vp] + [ctypes.c_int32] * 9 +
                                             [ctypes.c_double, vp, vp, vp,
                                              ctypes.POINTER(PhError)])
    L.ph_correction_input_device.restype = ctypes.c_int
    L.ph_sweep_config.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_uint32),
                                  ctypes.POINTER(ctypes.c_uint32)]
    L.ph_sweep_config.restype = None
    L.ph_gauss_noise_device.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, vp,
                                        ctypes.POINTER(PhError)]
    L.ph_gauss_noise_device.restype = ctypes.c_int
    L.ph_poisson_solve_batch_device.argtypes = [ctypes.POINTER(PhSolveDesc),
                                                ctypes.POINTER(PhBatchItem), ctypes.c_uint32, vp,
                                                ctypes.POINTER(PhReport), ctypes.POINTER(PhError)]
    L.ph_poisson_solve_batch_device.restype = ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    pe = ctypes.POINTER(PhError)
    L.ph_unet_param_storage.restype = ctypes.c_uint64
    L.ph_unet_param_count.restype = ctypes.c_uint64
    L.ph_unet_create.argtypes = [ctypes.c_int32, ctypes.c_int32, fp, ctypes.c_uint64,
                                 ctypes.POINTER(vp), pe]
    L.ph_unet_create.restype = ctypes.c_int
    L.ph_unet_destroy.argtypes = [vp]
    L.ph_unet_destroy.restype = None
    L.ph_unet_forward.argtypes = [vp, vp, vp, ctypes.c_int32, vp, pe]
    L.ph_unet_forward.restype = ctypes.c_int
    L.ph_unet_train_step.argtypes = [vp, vp, vp, vp, ctypes.c_int32, ctypes.POINTER(PhTrainOpts),
                                     ctypes.POINTER(ctypes.c_double), vp, pe]
    L.ph_unet_train_step.restype = ctypes.c_int
    L.ph_unet_get_state.argtypes = [vp, fp, fp, fp, fp, ctypes.POINTER(ctypes.c_int32), pe]
    L.ph_unet_get_state.restype = ctypes.c_int
    L.ph_unet_set_state.argtypes = [vp, fp, fp, fp, ctypes.c_int32, pe]
    L.ph_unet_set_state.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int32)
    L.ph_unet_gather_device.argtypes = [vp, vp, vp, ctypes.c_uint64, ip, ip, ctypes.c_int32,
                                        ctypes.c_int32, vp, vp, vp, vp, pe]
    L.ph_unet_gather_device.restype = ctypes.c_int
    L.ph_warm_start_set_weights.argtypes = [fp, ctypes.c_uint64, pe]
    L.ph_warm_start_set_weights.restype = ctypes.c_int
    L.ph_warm_coarsen_device.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, vp, vp, vp, vp, pe]
    L.ph_warm_coarsen_device.restype = ctypes.c_int
    _lib = L
    return L


def _check(status: int, err: PhError):
    if status != PH_OK:
        cls = int(err.error_class)
        name = ERROR_CLASSES[cls] if 0 <= cls < len(ERROR_CLASSES) else "execution"
        raise HipCallError(status, name, err.code.decode(), err.message.decode())


@dataclass
class SolveResult:
    iterations: int
    converged: bool
    stop_reason: str
    tolerance: float
    final_residual: float
    residual_history: list = field(default_factory=list)
    worst_channel: int = 0
    refinements: int = 0
    launches: int = 0
    unknowns: int = 0
    unknown_sweeps: int = 0
    device_ms: float = 0.0
    omega: float = 0.0
    half_sweeps: int = 0
    warm_start_used: int = 0  # channels that kept the learned start
    warm_start_rel: float = 0.0
    sweep_ms: float = 0.0
    coarse_corrections: int = 0


def make_desc(width, height, channels, max_iterations, tolerance, lam=0.0, op=PH_OP_POISSON_BLEND,
              omega=0.0, half_sweeps=0, flags=0, tile_rows=0) -> PhSolveDesc:
    return PhSolveDesc(width=width, height=height, channels=channels,
                       max_iterations=max_iterations, tolerance=tolerance, lam=lam,
                       omega=omega, flags=flags, half_sweeps=half_sweeps, op=op,
                       tile_rows=tile_rows)


def _result(rep: PhReport, hist: np.ndarray | None) -> SolveResult:
    return SolveResult(
        iterations=int(rep.iterations),
        converged=bool(rep.converged),
        stop_reason=STOP_NAMES.get(int(rep.stop_reason), "unknown"),
        tolerance=float(rep.tolerance),
        final_residual=float(rep.final_residual),
        residual_history=[] if hist is None else hist[: rep.history_len].tolist(),
        worst_channel=int(rep.worst_channel),
        refinements=int(rep.refinements),
        launches=int(rep.launches),
        unknowns=int(rep.unknowns),
        unknown_sweeps=int(rep.unknown_sweeps),
        device_ms=float(rep.device_ms),
        omega=float(rep.omega_used),
        half_sweeps=int(rep.half_sweeps_used),
        warm_start_used=int(rep.warm_start_used),
        warm_start_rel=float(rep.warm_start_rel),
        sweep_ms=float(rep.sweep_ms),
        coarse_corrections=int(rep.coarse_corrections),
    )


def abi_version() -> int:
    return int(lib().ph_abi_version())


def probe() -> dict:
    info, err = PhDeviceInfo(), PhError()
    st = lib().ph_probe(ctypes.byref(info), ctypes.byref(err))
    _check(st, err)
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "compute_units": int(info.compute_units),
        "device_index": int(info.device_index),
        "runtime_version": int(info.runtime_version),
        "driver_version": int(info.driver_version),
        "hbm_bytes": int(info.hbm_bytes),
    }


def force_unavailable(forced: bool) -> None:
    lib().ph_force_unavailable(1 if forced else 0)


def poisson_solve(desc: PhSolveDesc, source: np.ndarray, target: np.ndarray, mask: np.ndarray):
    """Host-buffer solve: returns (candidate f32 array shaped like target, SolveResult)."""
    source = np.ascontiguousarray(source, dtype=np.float32)
    target = np.ascontiguousarray(target, dtype=np.float32)
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    out = np.empty_like(target)
    cap = int(desc.max_iterations)
    hist = np.zeros(max(cap, 1), dtype=np.float64)
    rep = PhReport()
    rep.residual_history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rep.history_capacity = cap
    err = PhError()
    st = lib().ph_poisson_solve(ctypes.byref(desc), source.ctypes.data, target.ctypes.data,
                                mask.ctypes.data, out.ctypes.data, ctypes.byref(rep),
                                ctypes.byref(err))
    _check(st, err)
    return out, _result(rep, hist)


def poisson_solve_device(desc: PhSolveDesc, d_source: int, d_target: int, d_mask: int,
                         d_out: int, stream: int = 0, want_history: bool = True,
                         d_start: int | None = None) -> SolveResult:
    """Device-buffer solve on raw device pointers. With ``d_start`` the solve
    races that initial field against the anchor start
    (ph_poisson_solve_start_device)."""
    cap = int(desc.max_iterations)
    hist = np.zeros(max(cap, 1), dtype=np.float64)
    rep = PhReport()
    if want_history:
        rep.residual_history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep.history_capacity = cap
    err = PhError()
    if d_start is None:
        st = lib().ph_poisson_solve_device(ctypes.byref(desc), d_source, d_target, d_mask, d_out,
                                           stream or None, ctypes.byref(rep), ctypes.byref(err))
    else:
        st = lib().ph_poisson_solve_start_device(ctypes.byref(desc), d_source, d_target, d_mask,
                                                 d_start, d_out, stream or None,
                                                 ctypes.byref(rep), ctypes.byref(err))
    _check(st, err)
    return _result(rep, hist)


NOALLOC_SCRATCH_BYTES = 128 << 20  # PH_NOALLOC_SCRATCH_BYTES


def poisson_solve_device_noalloc(desc: PhSolveDesc, d_source: int, d_target: int, d_mask: int,
                                 d_out: int, stream: int = 0,
                                 want_history: bool = True) -> SolveResult:
    """ph_poisson_solve_device_noalloc: the plain device solve with no
    allocation in the call (its workspace is a fixed device array of
    NOALLOC_SCRATCH_BYTES)."""
    cap = int(desc.max_iterations)
    hist = np.zeros(max(cap, 1), dtype=np.float64)
    rep = PhReport()
    if want_history:
        rep.residual_history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rep.history_capacity = cap
    err = PhError()
    st = lib().ph_poisson_solve_device_noalloc(ctypes.byref(desc), d_source, d_target, d_mask,
                                               d_out, stream or None, ctypes.byref(rep),
                                               ctypes.byref(err))
    _check(st, err)
    return _result(rep, hist)


def fbm_device(d_out: int, width, height, frequency, seed, octaves, lacunarity=2.0, gain=0.5,
               channel_stride=1, channel_offset=0, scale=1.0, bias=0.0, stream: int = 0):
    err = PhError()
    st = lib().ph_fbm_device(d_out, width, height, channel_stride, channel_offset, frequency,
                             seed & 0xFFFFFFFFFFFFFFFF, octaves, lacunarity, gain, scale, bias,
                             stream or None, ctypes.byref(err))
    _check(st, err)


def polygon_mask_device(d_out: int, width, height, points, even_odd=False, stream: int = 0):
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1))
    err = PhError()
    st = lib().ph_polygon_mask_device(d_out, width, height,
                                      pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      len(pts) // 2, 1 if even_odd else 0, stream or None,
                                      ctypes.byref(err))
    _check(st, err)


def coarsen_plane_device(d_g: int, d_cls: int, pitch: int, pw: int, ph: int, ox: int, oy: int,
                         x0: int, y0: int, factor: int, res: int, d_in: int, stream: int = 0):
    """The solver's coarse-correction input from a plane (ph_coarsen_plane_device)."""
    err = PhError()
    st = lib().ph_coarsen_plane_device(d_g, d_cls, pitch, pw, ph, ox, oy, x0, y0, factor, res,
                                       d_in, stream or None, ctypes.byref(err))
    _check(st, err)


def sweep_config(work: float) -> tuple[int, int]:
    """(half-sweeps per launch, tile rows) a call of this work takes by default
    (ph_sweep_config): work = sum over its images of channels x the interior's
    bounding-box area."""
    k, hb = ctypes.c_uint32(0), ctypes.c_uint32(0)
    lib().ph_sweep_config(float(work), ctypes.byref(k), ctypes.byref(hb))
    return int(k.value), int(hb.value)


def correction_input_device(d_e: int, d_g: int, d_cls: int, pitch: int, pw: int, ph: int,
                            ox: int, oy: int, x0: int, y0: int, factor: int, res: int,
                            lam: float, d_in: int, d_rsq: int, stream: int = 0):
    """The correction-round input from a plane's state x + e
    (ph_correction_input_device)."""
    err = PhError()
    st = lib().ph_correction_input_device(d_e, d_g, d_cls, pitch, pw, ph, ox, oy, x0, y0, factor,
                                          res, lam, d_in, d_rsq, stream or None,
                                          ctypes.byref(err))
    _check(st, err)


def assert_no_change_outside_mask_device(d_before: int, d_after: int, d_mask: int, width: int,
                                         height: int, channels: int, decode_srgb: bool = False,
                                         outside_threshold: float = 0.0,
                                         coverage_epsilon: float = 0.0, stream: int = 0) -> dict:
    """assert.no_change_outside_mask@1's scan on device buffers
    (ph_assert_no_change_outside_mask_device). Returns the metrics with (x, y)
    coordinates."""
    r = PhOutsideScan()
    err = PhError()
    st = lib().ph_assert_no_change_outside_mask_device(
        d_before or None, d_after or None, d_mask or None, width, height, channels,
        int(bool(decode_srgb)), float(outside_threshold), float(coverage_epsilon),
        stream or None, ctypes.byref(r), ctypes.byref(err))
    _check(st, err)
    return dict(changed_outside=int(r.changed_outside), max_outside=float(r.max_outside),
                worst_pixel=None if r.worst_x < 0 else (int(r.worst_x), int(r.worst_y)),
                worst_delta=float(r.worst_delta),
                locations=[(int(r.locations[k][0]), int(r.locations[k][1]))
                           for k in range(r.n_locations)])


def ellipse_mask_device(d_out: int, width, height, cx, cy, rx, ry, stream: int = 0):
    err = PhError()
    st = lib().ph_ellipse_mask_device(d_out, width, height, cx, cy, rx, ry, stream or None,
                                      ctypes.byref(err))
    _check(st, err)


def gauss_noise_device(d_out: int, n: int, seed: int, sigma: float, stream: int = 0):
    err = PhError()
    st = lib().ph_gauss_noise_device(d_out, n, seed & 0xFFFFFFFFFFFFFFFF, sigma, stream or None,
                                     ctypes.byref(err))
    _check(st, err)


def poisson_solve_batch_device(desc: PhSolveDesc, items, stream: int = 0,
                               want_history: bool = True):
    """Batched solve. items: sequence of (width, height, channels, d_source,
    d_target, d_mask, d_out). Returns one SolveResult per item."""
    n = len(items)
    arr = (PhBatchItem * max(n, 1))()
    for i, (w, h, c, s_, t_, m_, o_) in enumerate(items):
        arr[i] = PhBatchItem(w, h, c, 0, s_, t_, m_, o_)
    reps = (PhReport * max(n, 1))()
    cap = int(desc.max_iterations)
    hists = [None] * n  # no history buffers unless asked for (max_iterations doubles each)
    if want_history:
        for i in range(n):
            hists[i] = np.zeros(max(cap, 1), dtype=np.float64)
            reps[i].residual_history = hists[i].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            reps[i].history_capacity = cap
    err = PhError()
    st = lib().ph_poisson_solve_batch_device(ctypes.byref(desc), arr, n, stream or None, reps,
                                             ctypes.byref(err))
    _check(st, err)
    return [_result(reps[i], hists[i]) for i in range(n)]


# ---------------------------------------------------------------------------
# learned warm start

def _fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def unet_param_storage() -> int:
    """FP32 values of the network's parameter storage (padding included)."""
    return int(lib().ph_unet_param_storage())


def unet_param_count() -> int:
    """Learnable values of the network (padding excluded)."""
    return int(lib().ph_unet_param_count())


class UNet:
    """The warm-start network with its training state (device-resident)."""

    def __init__(self, max_batch: int, res: int, params: np.ndarray | None = None, seed: int = 0):
        self.max_batch, self.res = int(max_batch), int(res)
        h = ctypes.c_void_p()
        err = PhError()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        st = lib().ph_unet_create(self.max_batch, self.res, _fptr(p), seed & 0xFFFFFFFFFFFFFFFF,
                                  ctypes.byref(h), ctypes.byref(err))
        _check(st, err)
        self._h = h

    def close(self):
        if self._h:
            lib().ph_unet_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, d_in: int, d_out: int, batch: int, stream: int = 0):
        err = PhError()
        _check(lib().ph_unet_forward(self._h, d_in, d_out, batch, stream or None,
                                     ctypes.byref(err)), err)

    def train_step(self, d_in: int, d_target: int, d_weight: int, batch: int, lr: float,
                   beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, update=True,
                   pyramid=0.0, stream: int = 0) -> float:
        o = PhTrainOpts(lr, beta1, beta2, eps, weight_decay, 1 if update else 0, pyramid)
        loss = ctypes.c_double(0.0)
        err = PhError()
        _check(lib().ph_unet_train_step(self._h, d_in, d_target, d_weight, batch, ctypes.byref(o),
                                        ctypes.byref(loss), stream or None, ctypes.byref(err)),
               err)
        return float(loss.value)

    def state(self):
        """(params, grads, adam_m, adam_v, step) as host arrays."""
        n = unet_param_storage()
        arrs = [np.zeros(n, np.float32) for _ in range(4)]
        step = ctypes.c_int32(0)
        err = PhError()
        _check(lib().ph_unet_get_state(self._h, *(_fptr(a) for a in arrs), ctypes.byref(step),
                                       ctypes.byref(err)), err)
        return (*arrs, int(step.value))

    def set_state(self, params, adam_m=None, adam_v=None, step: int = 0):
        err = PhError()
        conv = [None if a is None else np.ascontiguousarray(a, dtype=np.float32)
                for a in (params, adam_m, adam_v)]
        _check(lib().ph_unet_set_state(self._h, *(_fptr(a) for a in conv), int(step),
                                       ctypes.byref(err)), err)


def warm_start_set_weights(params: np.ndarray | None) -> None:
    """Set (or clear, with None) the weights of the solver's learned warm start."""
    err = PhError()
    if params is None:
        st = lib().ph_warm_start_set_weights(None, 0, ctypes.byref(err))
    else:
        p = np.ascontiguousarray(params, dtype=np.float32)
        st = lib().ph_warm_start_set_weights(_fptr(p), p.size, ctypes.byref(err))
    _check(st, err)


def warm_coarsen_device(d_source, d_anchor, d_mask, d_solution, width, height, channels, channel,
                        lam, x0, y0, factor, res, d_in, d_target=None, d_weight=None,
                        stream: int = 0, mode: int = 0):
    """Coarse net input (and target) of one channel; mode 0 = warm start,
    mode 1 = coarse-correction problem (P = 0, rhs = Laplacian of the source)."""
    err = PhError()
    st = lib().ph_warm_coarsen_device(d_source, d_anchor, d_mask, d_solution or None, width,
                                      height, channels, channel, lam, x0, y0, factor, res, mode,
                                      d_in, d_target or None, d_weight or None, stream or None,
                                      ctypes.byref(err))
    _check(st, err)


def unet_gather_device(d_in_all: int, d_target_all: int, d_weight_all: int, n_samples: int,
                       index, mode, res: int, d_in: int, d_target: int, d_weight: int,
                       stream: int = 0):
    """Gather a training batch (example index[j] under dihedral map mode[j])."""
    idx = np.ascontiguousarray(index, dtype=np.int32)
    md = None if mode is None else np.ascontiguousarray(mode, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int32)
    err = PhError()
    st = lib().ph_unet_gather_device(d_in_all, d_target_all, d_weight_all, n_samples,
                                     idx.ctypes.data_as(ip),
                                     None if md is None else md.ctypes.data_as(ip), idx.size,
                                     res, d_in, d_target, d_weight, stream or None,
                                     ctypes.byref(err))
    _check(st, err)
