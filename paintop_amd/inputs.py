"""Seeded synthetic workloads for the Poisson path.

Large images are generated on the device (``ph_fbm_device``,
``ph_polygon_mask_device``), so a 4096x4096 RGBA pair never crosses PCIe.
Small host-side cases use numpy. Every random draw comes from a SplitMix64
stream whose finalizer is paintop's ``mix64`` (crates/paintop-cpu/src/noise.rs:72-76).

Workloads:
  gaussian_bump_case   256x256 Gray: Gaussian-bump target, Laplacian guidance of a
                       second bump field, centred 80x60 ellipse mask (omega 1.9, 1e-5).
  seamless_clone_case  4096x4096 RGBA fbm source/target (frequency 0.125, 6 octaves,
                       gain 0.5, mapped to [0, 1]), alpha 1, 12-vertex star mask
                       covering ~30% of the image; solved to 1e-6.
  screened_case        8192x8192 Gray screened Poisson, lambda 0.01, full image
                       minus a 64 px border; solved to 1e-7.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = 0xFFFFFFFFFFFFFFFF


def mix64(z: int) -> int:
    z &= _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        return mix64(self.state)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def uniform_array(self, n: int, lo: float, hi: float) -> np.ndarray:
        return np.array([self.uniform(lo, hi) for _ in range(n)], dtype=np.float64)


# --- host-side small cases -------------------------------------------------------


def bump_field(width: int, height: int, seed: int, bumps: int = 8) -> np.ndarray:
    """Seeded sum of isotropic Gaussian bumps (sigma 10-40 px, amplitude
    U(0.2, 0.8)) plus U(0, 0.02) noise, clamped to [0, 1]; f32 (H, W)."""
    rng = SplitMix64(seed)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    x += 0.5
    y += 0.5
    f = np.zeros((height, width), dtype=np.float64)
    for _ in range(bumps):
        cx, cy = rng.uniform(0, width), rng.uniform(0, height)
        sig = rng.uniform(10.0, 40.0)
        amp = rng.uniform(0.2, 0.8)
        f += amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sig * sig))
    noise = rng.uniform_array(width * height, 0.0, 0.02).reshape(height, width)
    return np.clip(f + noise, 0.0, 1.0).astype(np.float32)


def ellipse_mask(width: int, height: int, cx: float, cy: float, rx: float, ry: float):
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    d = ((x + 0.5 - cx) / rx) ** 2 + ((y + 0.5 - cy) / ry) ** 2
    return (d <= 1.0).astype(np.float32)


def gaussian_bump_case(size: int = 256):
    """The 256x256 Gray CPU-reference case: (source, target, mask) f32 (H, W).
    The op's guidance is the Laplacian of ``source`` (seed 2); the target and
    Dirichlet boundary come from ``target`` (seed 1)."""
    target = bump_field(size, size, seed=1)
    source = bump_field(size, size, seed=2)
    m = ellipse_mask(size, size, size / 2.0, size / 2.0, 80.0 * size / 256, 60.0 * size / 256)
    return source, target, m


def star_polygon(cx: float, cy: float, r_outer: float, inner_ratio: float = 0.5,
                 points: int = 6, jitter: float = 0.1, seed: int = 301):
    """A 2*points-vertex star, radii alternating r_outer / inner_ratio*r_outer,
    each jittered radially by a seeded U(-jitter, +jitter) factor."""
    rng = SplitMix64(seed)
    verts = []
    for k in range(2 * points):
        r = r_outer if k % 2 == 0 else r_outer * inner_ratio
        r *= 1.0 + rng.uniform(-jitter, jitter)
        a = math.pi * k / points - math.pi / 2
        verts.append((cx + r * math.cos(a), cy + r * math.sin(a)))
    return verts


def random_polygon(cx, cy, rx, ry, n, rng: SplitMix64):
    """An n-vertex star-shaped polygon inside the ellipse (rx, ry)."""
    verts = []
    for k in range(n):
        a = 2 * math.pi * (k + rng.uniform(-0.3, 0.3)) / n
        s = rng.uniform(0.55, 1.0)
        verts.append((cx + s * rx * math.cos(a), cy + s * ry * math.sin(a)))
    return verts


# --- device-side large cases -----------------------------------------------------


def seamless_clone_case(size: int = 4096, channels: int = 4, star_seed: int = 301):
    """4096x4096 RGBA fbm blend generated in HBM. Returns DeviceArrays
    (source, target, mask): source/target (H, W, C), mask (H, W)."""
    from . import hip
    from .device import DeviceArray, synchronize

    src = DeviceArray((size, size, channels), np.float32)
    tgt = DeviceArray((size, size, channels), np.float32)
    msk = DeviceArray((size, size), np.float32)
    for c in range(channels):
        if c < 3:
            hip.fbm_device(src.ptr, size, size, 0.125, 101 + c, 6, 2.0, 0.5,
                           channel_stride=channels, channel_offset=c, scale=0.5, bias=0.5)
            hip.fbm_device(tgt.ptr, size, size, 0.125, 201 + c, 6, 2.0, 0.5,
                           channel_stride=channels, channel_offset=c, scale=0.5, bias=0.5)
        else:  # alpha = 1 (fmaf(v, 0, 1))
            for arr in (src, tgt):
                hip.fbm_device(arr.ptr, size, size, 1.0, 0, 1, 2.0, 0.5,
                               channel_stride=channels, channel_offset=c, scale=0.0, bias=1.0)
    r_outer = math.sqrt(0.30 * size * size / 1.5)
    verts = star_polygon(size / 2.0, size / 2.0, r_outer, 0.5, 6, 0.1, star_seed)
    hip.polygon_mask_device(msk.ptr, size, size, verts)
    synchronize()
    return src, tgt, msk


def screened_case(size: int = 8192, border: int = 64):
    """8192x8192 Gray screened case: anchor = (fbm+1)/2 (seed 401), guidance =
    (fbm2+1)/2 + N(0, 0.01) (seeds 402, 403), mask = 1 except a 64 px border.
    Returns DeviceArrays (guidance, anchor, mask)."""
    from . import hip
    from .device import DeviceArray, synchronize

    anchor = DeviceArray((size, size), np.float32)
    guide = DeviceArray((size, size), np.float32)
    msk = DeviceArray((size, size), np.float32)
    hip.fbm_device(anchor.ptr, size, size, 0.125, 401, 6, 2.0, 0.5, scale=0.5, bias=0.5)
    hip.fbm_device(guide.ptr, size, size, 0.125, 402, 6, 2.0, 0.5, scale=0.5, bias=0.5)
    hip.gauss_noise_device(guide.ptr, size * size, 403, 0.01)
    b = float(border)
    hip.polygon_mask_device(msk.ptr, size, size,
                            [(b, b), (size - b, b), (size - b, size - b), (b, size - b)])
    synchronize()
    return guide, anchor, msk


def batch_case(n: int = 512, seed0: int = 500):
    """512 independent single-channel blends of uneven size: side s = exp U(ln 32,
    ln 1024), aspect U(0.5, 2), W = round(s sqrt(a)), H = round(s / sqrt(a)); the
    mask is a centred ellipse or a 6-10 vertex polygon filling U(0.2, 0.6) of the
    box; source and target are fbm (frequency 0.5, 6 octaves) mapped to [0, 1].
    Instance i draws from SplitMix64(seed0 + i). Returns a list of
    (width, height, source, target, mask) DeviceArrays."""
    from . import hip
    from .device import DeviceArray, synchronize

    out = []
    for i in range(n):
        rng = SplitMix64(seed0 + i)
        s = math.exp(rng.uniform(math.log(32.0), math.log(1024.0)))
        a = rng.uniform(0.5, 2.0)
        w = max(3, int(round(s * math.sqrt(a))))
        h = max(3, int(round(s / math.sqrt(a))))
        fill = rng.uniform(0.2, 0.6)
        src = DeviceArray((h, w), np.float32)
        tgt = DeviceArray((h, w), np.float32)
        msk = DeviceArray((h, w), np.float32)
        hip.fbm_device(src.ptr, w, h, 0.5, 10_000 + 2 * i, 6, 2.0, 0.5, scale=0.5, bias=0.5)
        hip.fbm_device(tgt.ptr, w, h, 0.5, 10_001 + 2 * i, 6, 2.0, 0.5, scale=0.5, bias=0.5)
        if rng.uniform() < 0.5:
            k = math.sqrt(4.0 * fill / math.pi)
            hip.ellipse_mask_device(msk.ptr, w, h, w / 2.0, h / 2.0, k * w / 2.0, k * h / 2.0)
        else:
            nv = 6 + int(rng.uniform() * 5) % 5
            verts = random_polygon(w / 2.0, h / 2.0, w / 2.0, h / 2.0, nv, rng)
            x = np.array([v[0] for v in verts]) - w / 2.0
            y = np.array([v[1] for v in verts]) - h / 2.0
            area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
            sc = math.sqrt(fill * w * h / max(area, 1e-9))
            verts = [(w / 2.0 + sc * xx, h / 2.0 + sc * yy) for xx, yy in zip(x, y)]
            hip.polygon_mask_device(msk.ptr, w, h, verts)
        out.append((w, h, src, tgt, msk))
    synchronize()
    return out


def spots_case(size: int = 4096, n_spots: int = 48, channels: int = 4, seed: int = 601):
    """Blemish-style blend: fbm source/target (as seamless_clone_case) and a mask
    of n_spots disjoint elliptical spots (semi-axes 12..96 px) scattered over
    the image. Returns DeviceArrays (source, target) and the host mask."""
    from . import hip
    from .device import DeviceArray, synchronize

    src = DeviceArray((size, size, channels), np.float32)
    tgt = DeviceArray((size, size, channels), np.float32)
    for c in range(channels):
        hip.fbm_device(src.ptr, size, size, 0.125, 101 + c, 6, 2.0, 0.5,
                       channel_stride=channels, channel_offset=c, scale=0.5, bias=0.5)
        hip.fbm_device(tgt.ptr, size, size, 0.125, 201 + c, 6, 2.0, 0.5,
                       channel_stride=channels, channel_offset=c, scale=0.5, bias=0.5)
    rng = SplitMix64(seed)
    msk = np.zeros((size, size), np.float32)
    placed = []
    tries = 0
    while len(placed) < n_spots and tries < 20 * n_spots:
        tries += 1
        rx, ry = rng.uniform(12.0, 96.0), rng.uniform(12.0, 96.0)
        cx, cy = rng.uniform(rx + 4, size - rx - 4), rng.uniform(ry + 4, size - ry - 4)
        if any(abs(cx - px) < rx + prx + 6 and abs(cy - py) < ry + pry + 6
               for px, py, prx, pry in placed):
            continue
        placed.append((cx, cy, rx, ry))
        y0, y1 = int(cy - ry) - 1, int(cy + ry) + 2
        x0, x1 = int(cx - rx) - 1, int(cx + rx) + 2
        yy, xx = np.mgrid[y0:y1, x0:x1]
        inside = ((xx + 0.5 - cx) / rx) ** 2 + ((yy + 0.5 - cy) / ry) ** 2 <= 1.0
        msk[y0:y1, x0:x1][inside] = 1.0
    synchronize()
    return src, tgt, msk
