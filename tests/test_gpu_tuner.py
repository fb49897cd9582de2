"""The SOR-factor tuner on the GPU, through the C ABI.

With no caller omega the solver sets omega = 2 / (1 + sqrt(1 - rho^2)),
rho = (4 - mu) / (4 + lambda), where mu is the smallest eigenvalue of the
5-point operator on the interior: measured on the coarse mask (a block of
f x f cells is inside when at least half of them are interior) and scaled by
1 / f^2 for domains of at least 32768 interior cells, else the larger of the
Faber-Krahn disc and bounding-rectangle bounds. The coarse mask is built on the
device and the eigenvalues are cached per coarse mask under a key the device
computes; these tests pin the factor against scipy's eigenvalue of a numpy
restatement of the coarse mask, and check that a cache hit, a cache miss and
a bypassed cache give the same factor and the same field.
"""
import math

import numpy as np
import pytest

from paintop_amd import hip
from tests.helpers import classify, smallest_eigenvalue

pytestmark = pytest.mark.gpu


def _region(cells):
    """The solver's region around the interior: its bounding box plus the
    ring, widened one column to the left when needed so that the region's
    origin has an even x + y."""
    ys, xs = np.nonzero(cells)
    rx0, ry0 = int(xs.min()) - 1, int(ys.min()) - 1
    rx1, ry1 = int(xs.max()) + 2, int(ys.max()) + 2
    if (rx0 + ry0) & 1:
        rx0 -= 1
    return rx0, ry0, rx1 - rx0, ry1 - ry0


def _omega(mu, lam=0.0):
    rho = (4.0 - min(mu, 4.0)) / (4.0 + lam)
    return min(max(2.0 / (1.0 + math.sqrt(max(1.0 - rho * rho, 0.0))), 1.0), 1.9999)


def _expected_omega(mask):
    cells = classify(mask)
    n = int(cells.sum())
    rx0, ry0, rw, rh = _region(cells)
    if n < 32768:
        j01 = 2.404825557695773
        r = math.sqrt(max(n, 1) / math.pi)
        mu = max((j01 / r) ** 2, (math.pi / rw) ** 2 + (math.pi / rh) ** 2)
        return _omega(min(mu, 4.0)), None
    f = max(1, (max(rw, rh) + 127) // 128)
    cw, ch = -(-rw // f), -(-rh // f)
    h, w = cells.shape
    # the region in coarse blocks; cells outside the image count as exterior
    pad = np.zeros((ch * f, cw * f), np.int64)
    y0, x0 = max(ry0, 0), max(rx0, 0)
    y1, x1 = min(ry0 + ch * f, h), min(rx0 + cw * f, w)
    pad[y0 - ry0:y1 - ry0, x0 - rx0:x1 - rx0] = cells[y0:y1, x0:x1]
    counts = pad.reshape(ch, f, cw, f).sum(axis=(1, 3))
    inside = (2 * counts >= f * f).astype(np.uint8)
    mu_c = smallest_eigenvalue(np.pad(inside, 1))
    return _omega(mu_c / (f * f)), (f, cw, ch)


def _solve(src, tgt, msk, n, c, iters=40):
    from paintop_amd.device import DeviceArray

    out = DeviceArray(tgt.shape, np.float32)
    d = hip.make_desc(n, n, c, iters, 1e-12)
    r = hip.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, out.ptr)
    return r, out.numpy()


@pytest.mark.parametrize("n,seed", [(640, 301), (900, 302)])
def test_tuned_omega_follows_the_coarse_mask_eigenvalue(gpu, n, seed):
    """The factor from the device-built coarse mask and the host eigen solve
    (inverse iteration with CG) equals the one from scipy's eigenvalue of the
    numpy coarse mask within 2e-5 (about 1e-3 of mu at these sizes)."""
    from paintop_amd.inputs import seamless_clone_case

    src, tgt, msk = seamless_clone_case(n, channels=1, star_seed=seed)
    want, geom = _expected_omega(msk.numpy())
    assert geom is not None  # the precise tuner runs at these sizes
    r, _ = _solve(src, tgt, msk, n, 1)
    print(f"[tuner] {n}^2 seed {seed}: f, cw, ch = {geom}; omega gpu {r.omega:.7f} "
          f"scipy {want:.7f}", flush=True)
    assert abs(r.omega - want) <= 2e-5


def test_small_domains_use_the_geometric_bound(gpu):
    from paintop_amd.inputs import seamless_clone_case

    n = 200  # a ~30% star: about 12 000 interior cells
    src, tgt, msk = seamless_clone_case(n, channels=1)
    want, geom = _expected_omega(msk.numpy())
    assert geom is None
    r, _ = _solve(src, tgt, msk, n, 1)
    assert abs(r.omega - want) <= 1e-12


def test_cache_hit_miss_and_bypass_agree(gpu, monkeypatch):
    """A mask seen for the first time (its coarse mask copied from the device
    and solved), the same mask again (a hit on the device-computed key) and a
    bypassed cache give the same factor and the same field, bit for bit; a
    different mask gets its own factor."""
    from paintop_amd.inputs import seamless_clone_case

    n = 700
    src, tgt, msk = seamless_clone_case(n, channels=3, star_seed=777)
    r1, f1 = _solve(src, tgt, msk, n, 3)
    r2, f2 = _solve(src, tgt, msk, n, 3)
    monkeypatch.setenv("PAINTOP_HIP_TUNE_CACHE", "0")
    r3, f3 = _solve(src, tgt, msk, n, 3)
    monkeypatch.delenv("PAINTOP_HIP_TUNE_CACHE")
    assert r1.omega == r2.omega == r3.omega
    assert f1.tobytes() == f2.tobytes() == f3.tobytes()
    s2, t2, m2 = seamless_clone_case(n, channels=3, star_seed=778)
    r4, _ = _solve(s2, t2, m2, n, 3)
    want4, _ = _expected_omega(m2.numpy())
    assert r4.omega != r1.omega and abs(r4.omega - want4) <= 2e-5
