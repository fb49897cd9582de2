"""Write the full-size parity fixtures: an independent FP64 solution of the
4096^2 seamless-clone blend (R, G, B) and of the 8192^2 screened-Poisson case,
each driven to a 1e-12 relative residual (or the FP64 floor of the residual
evaluation, at most 2e-12) on the CPU by the oracle's
multigrid-preconditioned CG (oracle/tight_solve.c), which is pinned against a
sparse direct solve in tests/test_oracle.py.

The inputs are rebuilt on the CPU with the oracle's generators, which match
the device generators the GPU tests use: fbm bit for bit (oracle/generators.c
restates noise.rs), the polygon mask bit for bit (the same ray cast), the
Gaussian noise within rounding of the host's log/cos. The fixture records
input samples so a GPU test can confirm it solved the same system.

Per solved channel it stores:
  pts      4096 seeded interior cells (flat row-major indices)
  u_pts    the FP64 solution there
  src_pts, tgt_pts  the f32 inputs there (source/guidance, target/anchor)
  blk_mean, blk_n   the mean of the solution over the interior cells of each
           64x64-pixel (4096^2) / 128x128-pixel (8192^2) block, and their count
  r0, final_residual, iterations (the solve's own report)
and per image the interior-cell count and a CRC32 of the packed classification.

Run: python tests/golden/make_full_size_fixtures.py [--only seamless|screened]
(about 1-3 minutes per case on 8 cores; needs ~10 GB of memory for 8192^2).
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from paintop_amd.inputs import SplitMix64, star_polygon  # noqa: E402
from tests.helpers import classify, laplacian  # noqa: E402

N_POINTS = 4096


def fbm01(n, seed):
    f = orc.fbm(n, n, 0.125, seed, 6)
    return f * np.float32(0.5) + np.float32(0.5)  # fmaf(v, 0.5, 0.5): the product is exact


def sample_points(cells: np.ndarray, seed: int, k: int = N_POINTS) -> np.ndarray:
    idx = np.flatnonzero(cells.reshape(-1))
    rng = SplitMix64(seed)
    pick = np.array([rng.next_u64() % idx.size for _ in range(k)], dtype=np.int64)
    return np.sort(idx[pick])


def block_means(u: np.ndarray, cells: np.ndarray, b: int):
    h, w = cells.shape
    m = cells.astype(np.float64)
    s = (u * m).reshape(h // b, b, w // b, b).sum(axis=(1, 3))
    c = m.reshape(h // b, b, w // b, b).sum(axis=(1, 3))
    return np.where(c > 0, s / np.maximum(c, 1), 0.0), c.astype(np.int64)


def solve_channel(cells, src, tgt, lam, tag, out, blk):
    rhs = np.where(cells == 1, laplacian(src), 0.0)
    t0 = time.time()
    u, rep = orc.tight_solve(cells, rhs, tgt.astype(np.float64), lam, tolerance=1e-12,
                             alpha=1.8)
    print(f"[{tag}] tight solve: {rep} in {time.time() - t0:.1f} s", flush=True)
    # 1e-12, or the FP64 floor of the residual evaluation just above it (the
    # field's own rounding, ~0.5 ulp per cell, alone gives ~6e-13 relative on
    # the 4096^2 channels, whose r0 is ~0.5): a restart that no longer halves
    # the true residual ends the solve
    assert rep["final_residual"] <= 2e-12, rep
    pts = out["pts"]
    out[f"{tag}_u_pts"] = u.reshape(-1)[pts]
    out[f"{tag}_src_pts"] = src.reshape(-1)[pts]
    out[f"{tag}_tgt_pts"] = tgt.reshape(-1)[pts]
    mean, cnt = block_means(u, cells, blk)
    out[f"{tag}_blk_mean"] = mean
    out[f"{tag}_blk_n"] = cnt
    out[f"{tag}_report"] = np.array([rep["r0"], rep["final_residual"], rep["iterations"],
                                     rep["restarts"]], dtype=np.float64)


def seamless(path):
    n = 4096
    r_outer = math.sqrt(0.30 * n * n / 1.5)
    verts = star_polygon(n / 2.0, n / 2.0, r_outer, 0.5, 6, 0.1, 301)
    m = orc.polygon_mask(n, n, verts)
    cells = classify(m)
    out = {"pts": sample_points(cells, 7001), "n_interior": np.int64(cells.sum()),
           "cells_crc32": np.int64(zlib.crc32(np.packbits(cells).tobytes())),
           "size": np.int64(n), "lam": np.float64(0.0), "block": np.int64(64)}
    for c in range(3):
        src = fbm01(n, 101 + c)
        tgt = fbm01(n, 201 + c)
        solve_channel(cells, src, tgt, 0.0, f"c{c}", out, 64)
    np.savez_compressed(path, **out)


def screened(path):
    n, border = 8192, 64
    anchor = fbm01(n, 401)
    guide = fbm01(n, 402)
    orc.gauss_noise(guide, 403, 0.01)
    b = float(border)
    m = orc.polygon_mask(n, n, [(b, b), (n - b, b), (n - b, n - b), (b, n - b)])
    cells = classify(m)
    del m
    out = {"pts": sample_points(cells, 7002), "n_interior": np.int64(cells.sum()),
           "cells_crc32": np.int64(zlib.crc32(np.packbits(cells).tobytes())),
           "size": np.int64(n), "lam": np.float64(0.01), "block": np.int64(128)}
    solve_channel(cells, guide, anchor, 0.01, "c0", out, 128)
    np.savez_compressed(path, **out)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["seamless", "screened"])
    a = ap.parse_args()
    if a.only in (None, "seamless"):
        seamless(os.path.join(HERE, "seamless_clone_4096_tight.npz"))
    if a.only in (None, "screened"):
        screened(os.path.join(HERE, "screened_8192_tight.npz"))
