"""ctypes face of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the tests, ``__graft_entry__.smoke()`` and the
benchmark's ``cpu_baseline`` leg import this module as the checker. The
product package never imports it.

Each function names the reference code it restates (bobisme/paintop):
``solve`` -> crates/paintop-cpu/src/poisson.rs:168-265,
``blend_channels`` -> crates/paintop-cpu/src/poisson_blend.rs:168-246,
``fbm`` -> crates/paintop-cpu/src/noise.rs:72-160, 218-240, 660-694,
``outside_scan`` -> crates/paintop-cpu/src/assert.rs:72-85, 178-201, 405-446.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

STOP_NAMES = {0: "converged", 1: "max-iterations", 2: "stalled"}


class _Report(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("stop_reason", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("_pad", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("final_residual", ctypes.c_double),
        ("history_len", ctypes.c_uint32),
        ("_pad2", ctypes.c_uint32),
    ]


class _TightReport(ctypes.Structure):
    _fields_ = [
        ("r0", ctypes.c_double),
        ("final_residual", ctypes.c_double),
        ("iterations", ctypes.c_uint32),
        ("restarts", ctypes.c_uint32),
        ("levels", ctypes.c_uint32),
        ("converged", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("_pad", ctypes.c_int32),
    ]


class _OutsideResult(ctypes.Structure):
    _fields_ = [
        ("changed_outside", ctypes.c_uint64),
        ("max_outside", ctypes.c_double),
        ("worst_index", ctypes.c_int64),
        ("worst_delta", ctypes.c_double),
        ("n_locations", ctypes.c_uint32),
        ("_pad", ctypes.c_uint32),
        ("locations", ctypes.c_int64 * 256),
    ]


@dataclass
class SolveReport:
    """Mirror of ``SolveReport`` (poisson.rs:126-141)."""

    iterations: int
    converged: bool
    stop_reason: str
    tolerance: float
    final_residual: float
    residual_history: list = field(default_factory=list)


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        up = ctypes.POINTER(ctypes.c_uint8)
        sz = ctypes.c_size_t
        L.or_classify_cells.argtypes = [fp, sz, sz, up]
        L.or_guidance_laplacian.argtypes = [dp, sz, sz, sz, sz]
        L.or_guidance_laplacian.restype = ctypes.c_double
        L.or_residual_l2.argtypes = [sz, sz, up, dp, dp, ctypes.c_double, dp]
        L.or_residual_l2.restype = ctypes.c_double
        L.or_solve.argtypes = [sz, sz, up, dp, dp, ctypes.c_double, ctypes.c_uint32,
                               ctypes.c_double, ctypes.c_double, dp, dp,
                               ctypes.POINTER(_Report)]
        L.or_blend_channels.argtypes = [fp, fp, fp, sz, sz, sz, ctypes.c_uint32,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        fp, dp, ctypes.POINTER(_Report)]
        L.or_mix64.argtypes = [ctypes.c_uint64]
        L.or_mix64.restype = ctypes.c_uint64
        L.or_lattice_value.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64]
        L.or_lattice_value.restype = ctypes.c_double
        L.or_value_noise.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        L.or_value_noise.restype = ctypes.c_double
        L.or_fbm_eval.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                  ctypes.c_uint32, ctypes.c_double, ctypes.c_double]
        L.or_fbm_eval.restype = ctypes.c_double
        L.or_fbm.argtypes = [sz, sz, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                             ctypes.c_double, ctypes.c_double, fp]
        L.or_rbsor_omp.argtypes = [sz, sz, up, dp, dp, ctypes.c_double, ctypes.c_uint32,
                                   ctypes.c_double, ctypes.c_double, dp, ctypes.c_int,
                                   ctypes.c_uint32, dp, dp, ctypes.POINTER(_Report)]
        L.or_rbsor_omp.restype = ctypes.c_int
        L.or_tight_solve.argtypes = [sz, sz, up, dp, dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_uint32, ctypes.c_double, ctypes.c_int, dp,
                                     ctypes.POINTER(_TightReport)]
        L.or_tight_solve.restype = ctypes.c_int
        L.or_polygon_mask.argtypes = [sz, sz, dp, sz, ctypes.c_int, fp]
        L.or_gauss_noise.argtypes = [fp, sz, ctypes.c_uint64, ctypes.c_double]
        L.or_outside_scan.argtypes = [fp, fp, fp, sz, sz, sz, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.POINTER(_OutsideResult)]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _report(r: _Report, hist: np.ndarray) -> SolveReport:
    return SolveReport(
        iterations=int(r.iterations),
        converged=bool(r.converged),
        stop_reason=STOP_NAMES[int(r.stop_reason)],
        tolerance=float(r.tolerance),
        final_residual=float(r.final_residual),
        residual_history=[float(v) for v in hist[: r.history_len]],
    )


def classify_cells(mask: np.ndarray) -> np.ndarray:
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    h, w = mask.shape
    cells = np.zeros((h, w), dtype=np.uint8)
    lib().or_classify_cells(_p(mask, ctypes.c_float), w, h, _p(cells, ctypes.c_uint8))
    return cells


def guidance_laplacian(field_: np.ndarray) -> np.ndarray:
    """rhs = Laplacian of ``field_`` at every pixel (poisson.rs:364-373)."""
    f = np.ascontiguousarray(field_, dtype=np.float64)
    h, w = f.shape
    out = np.empty_like(f)
    L = lib()
    fp = _p(f, ctypes.c_double)
    for r in range(h):
        for c in range(w):
            out[r, c] = L.or_guidance_laplacian(fp, c, r, w, h)
    return out


def residual_l2(cells, rhs, anchor, lam, field_) -> float:
    cells = np.ascontiguousarray(cells, dtype=np.uint8)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    anchor = np.ascontiguousarray(anchor, dtype=np.float64)
    field_ = np.ascontiguousarray(field_, dtype=np.float64)
    h, w = cells.shape
    return float(lib().or_residual_l2(w, h, _p(cells, ctypes.c_uint8), _p(rhs, ctypes.c_double),
                                      _p(anchor, ctypes.c_double), float(lam),
                                      _p(field_, ctypes.c_double)))


def solve(cells, rhs, anchor, lam, max_iterations, tolerance, omega):
    """poisson::solve restated (lexicographic SOR, f64)."""
    cells = np.ascontiguousarray(cells, dtype=np.uint8)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    anchor = np.ascontiguousarray(anchor, dtype=np.float64)
    h, w = cells.shape
    out = np.empty((h, w), dtype=np.float64)
    hist = np.zeros(max(int(max_iterations), 1), dtype=np.float64)
    rep = _Report()
    lib().or_solve(w, h, _p(cells, ctypes.c_uint8), _p(rhs, ctypes.c_double),
                   _p(anchor, ctypes.c_double), float(lam), int(max_iterations),
                   float(tolerance), float(omega), _p(out, ctypes.c_double),
                   _p(hist, ctypes.c_double), ctypes.byref(rep))
    return out, _report(rep, hist)


def blend_channels(source, target, mask, lam, max_iterations, tolerance, omega=1.8):
    """poisson_blend::blend_channels restated. source/target: (H, W, C) f32."""
    source = np.ascontiguousarray(source, dtype=np.float32)
    target = np.ascontiguousarray(target, dtype=np.float32)
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    if source.ndim == 2:
        source = source[:, :, None]
        target = target[:, :, None]
    h, w, c = source.shape
    out = np.empty_like(target)
    hist = np.zeros(max(int(max_iterations), 1), dtype=np.float64)
    rep = _Report()
    lib().or_blend_channels(_p(source, ctypes.c_float), _p(target, ctypes.c_float),
                            _p(mask, ctypes.c_float), c, w, h, int(max_iterations),
                            float(tolerance), float(lam), float(omega),
                            _p(out, ctypes.c_float), _p(hist, ctypes.c_double),
                            ctypes.byref(rep))
    return out, _report(rep, hist)


def rbsor_omp(cells, rhs, anchor, lam, max_iterations, tolerance, omega, threads=0,
              stall_window=16, omega_tab=None):
    """The all-cores red-black SOR stand-in for cpu.optimized (stall_window 0
    disables the stall guard, for timing a fixed number of sweeps)."""
    cells = np.ascontiguousarray(cells, dtype=np.uint8)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    anchor = np.ascontiguousarray(anchor, dtype=np.float64)
    h, w = cells.shape
    out = np.empty((h, w), dtype=np.float64)
    hist = np.zeros(max(int(max_iterations), 1), dtype=np.float64)
    rep = _Report()
    tab = None
    if omega_tab is not None:
        tab = np.ascontiguousarray(omega_tab, dtype=np.float64)
        assert tab.size >= 2 * int(max_iterations)
    used = lib().or_rbsor_omp(w, h, _p(cells, ctypes.c_uint8), _p(rhs, ctypes.c_double),
                              _p(anchor, ctypes.c_double), float(lam), int(max_iterations),
                              float(tolerance), float(omega),
                              None if tab is None else _p(tab, ctypes.c_double), int(threads),
                              int(stall_window), _p(out, ctypes.c_double),
                              _p(hist, ctypes.c_double),
                              ctypes.byref(rep))
    return out, _report(rep, hist), int(used)


def tight_solve(cells, rhs, anchor, lam, tolerance=1e-12, max_iterations=20000, alpha=1.0,
                threads=0):
    """The system of poisson.rs:168-240 solved to a tight FP64 relative residual
    by multigrid-preconditioned CG (oracle/tight_solve.c). Returns (field,
    report dict): r0, final_residual (the reference's residual formula and
    order), CG iterations, restarts, multigrid levels, threads."""
    cells = np.ascontiguousarray(cells, dtype=np.uint8)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    anchor = np.ascontiguousarray(anchor, dtype=np.float64)
    h, w = cells.shape
    out = np.empty((h, w), dtype=np.float64)
    rep = _TightReport()
    lib().or_tight_solve(w, h, _p(cells, ctypes.c_uint8), _p(rhs, ctypes.c_double),
                         _p(anchor, ctypes.c_double), float(lam), float(tolerance),
                         int(max_iterations), float(alpha), int(threads),
                         _p(out, ctypes.c_double), ctypes.byref(rep))
    return out, dict(r0=rep.r0, final_residual=rep.final_residual, iterations=rep.iterations,
                     restarts=rep.restarts, levels=rep.levels, converged=bool(rep.converged),
                     threads=rep.threads)


def polygon_mask(width, height, points, even_odd=False) -> np.ndarray:
    """mask.polygon@1's inside test at pixel centres (mask_polygon.rs:161-195)."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1))
    out = np.empty((height, width), dtype=np.float32)
    lib().or_polygon_mask(width, height, _p(pts, ctypes.c_double), pts.size // 2,
                          int(bool(even_odd)), _p(out, ctypes.c_float))
    return out


def gauss_noise(out: np.ndarray, seed: int, sigma: float) -> None:
    """out += sigma * N(0, 1) in place (seeded SplitMix64 + Box-Muller)."""
    assert out.dtype == np.float32 and out.flags.c_contiguous
    lib().or_gauss_noise(_p(out, ctypes.c_float), out.size, seed & 0xFFFFFFFFFFFFFFFF,
                         float(sigma))


def mix64(z: int) -> int:
    return int(lib().or_mix64(z & 0xFFFFFFFFFFFFFFFF))


def lattice_value(gx: int, gy: int, seed: int) -> float:
    return float(lib().or_lattice_value(gx, gy, seed & 0xFFFFFFFFFFFFFFFF))


def value_noise(x: float, y: float, seed: int) -> float:
    return float(lib().or_value_noise(x, y, seed & 0xFFFFFFFFFFFFFFFF))


def fbm_eval(x, y, seed, octaves, lacunarity=2.0, gain=0.5) -> float:
    return float(lib().or_fbm_eval(x, y, seed & 0xFFFFFFFFFFFFFFFF, octaves, lacunarity, gain))


def fbm(width, height, frequency, seed, octaves, lacunarity=2.0, gain=0.5) -> np.ndarray:
    out = np.empty((height, width), dtype=np.float32)
    lib().or_fbm(width, height, float(frequency), seed & 0xFFFFFFFFFFFFFFFF, int(octaves),
                 float(lacunarity), float(gain), _p(out, ctypes.c_float))
    return out


def outside_scan(before: np.ndarray, after: np.ndarray, mask: np.ndarray, decode: bool = False,
                 outside_threshold: float = 0.0, coverage_epsilon: float = 0.0) -> dict:
    """assert.no_change_outside_mask@1's scan (assert.rs:405-446). before/after
    (H, W, C) or (H, W) f32, mask (H, W) f32. Returns the scan with (x, y)
    coordinates (assert.rs:196-201)."""
    b = np.ascontiguousarray(before, dtype=np.float32)
    a = np.ascontiguousarray(after, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.float32)
    h, w = m.shape
    c = 1 if b.ndim == 2 else b.shape[2]
    r = _OutsideResult()
    fp = ctypes.c_float
    lib().or_outside_scan(_p(b, fp), _p(a, fp), _p(m, fp), w, h, c, int(bool(decode)),
                          float(outside_threshold), float(coverage_epsilon), ctypes.byref(r))
    xy = lambda i: (int(i) % w, int(i) // w)  # noqa: E731
    return dict(changed_outside=int(r.changed_outside), max_outside=float(r.max_outside),
                worst_pixel=None if r.worst_index < 0 else xy(r.worst_index),
                worst_delta=float(r.worst_delta),
                locations=[xy(r.locations[k]) for k in range(r.n_locations)])


def sor_spectral_radius(mu: float, lam: float, omega: float) -> float:
    """Spectral radius of lexicographic SOR (poisson.rs:244-265) on the 5-point
    system with smallest operator eigenvalue mu (of (4 + lam) I - N on the
    interior): the 5-point matrix in natural order is consistently ordered, so
    Young's relation (l + omega - 1)^2 = l omega^2 rho_J^2 holds with the
    Jacobi radius rho_J = (4 - mu) / (4 + lam). Below the optimal omega the
    largest root is ((omega rho_J + sqrt(omega^2 rho_J^2 - 4 (omega - 1))) / 2)^2;
    above it the radius is omega - 1."""
    import math

    rj = (4.0 - mu) / (4.0 + lam)
    disc = omega * omega * rj * rj - 4.0 * (omega - 1.0)
    if disc <= 0.0:
        return omega - 1.0
    return ((omega * rj + math.sqrt(disc)) / 2.0) ** 2


def project_sor_sweeps(mu: float, lam: float, omega: float, tolerance: float) -> float:
    """Sweeps lexicographic SOR needs for its relative residual to fall to
    `tolerance`, projected from its asymptotic rate (the spectral radius)."""
    import math

    rho = sor_spectral_radius(mu, lam, omega)
    return math.log(tolerance) / math.log(rho)


def mu_from_optimal_omega(omega_opt: float, lam: float = 0.0) -> float:
    """Invert omega_opt = 2 / (1 + sqrt(1 - rho_J^2)) for the smallest operator
    eigenvalue mu = 4 - (4 + lam) rho_J."""
    import math

    rj = math.sqrt(max(0.0, 1.0 - (2.0 / omega_opt - 1.0) ** 2))
    return 4.0 - (4.0 + lam) * rj
