"""Independent checks used by the tests: a sparse direct FP64 solve of the
5-point screened-Poisson system (a textbook LU via scipy), and the FP64
residual of a field. They pin the oracle and the GPU path to what the
mathematics fixes, independently of the iteration used."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def classify(mask: np.ndarray) -> np.ndarray:
    """Interior iff mask > 0.5 and not on the border (the closed form of
    crates/paintop-cpu/src/poisson.rs:340-355)."""
    h, w = mask.shape
    cells = np.zeros((h, w), dtype=np.uint8)
    if w >= 3 and h >= 3:
        cells[1:-1, 1:-1] = (mask[1:-1, 1:-1] > np.float32(0.5)).astype(np.uint8)
    return cells


def laplacian(f: np.ndarray) -> np.ndarray:
    """5-point Laplacian at interior positions (borders left 0)."""
    f = f.astype(np.float64)
    out = np.zeros_like(f)
    out[1:-1, 1:-1] = (f[1:-1, :-2] + f[1:-1, 2:] + f[:-2, 1:-1] + f[2:, 1:-1]) - 4.0 * f[1:-1, 1:-1]
    return out


def direct_solve(cells: np.ndarray, rhs: np.ndarray, anchor: np.ndarray, lam: float) -> np.ndarray:
    """Solve (Delta - lam) u = rhs - lam*anchor on interior cells, u = anchor
    elsewhere, exactly (sparse LU). Returns the full f64 field."""
    h, w = cells.shape
    idx = -np.ones((h, w), dtype=np.int64)
    ys, xs = np.nonzero(cells)
    n = len(ys)
    idx[ys, xs] = np.arange(n)
    u = anchor.astype(np.float64).copy()
    if n == 0:
        return u
    rows, cols, vals = [], [], []
    b = (rhs[ys, xs] - lam * anchor[ys, xs]).astype(np.float64)
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(np.full(n, -(4.0 + lam)))
    for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        ny, nx = ys + dy, xs + dx
        nb = idx[ny, nx]
        inner = nb >= 0
        rows.append(np.arange(n)[inner])
        cols.append(nb[inner])
        vals.append(np.ones(int(inner.sum())))
        b[~inner] -= anchor[ny[~inner], nx[~inner]]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    u[ys, xs] = spla.spsolve(A.tocsc(), b)
    return u


def residual_norm(cells, rhs, anchor, lam, field) -> float:
    """||(Delta - lam) u - (rhs - lam*anchor)||_2 over interior cells, in f64."""
    f = field.astype(np.float64)
    nsum = np.zeros_like(f)
    nsum[1:-1, 1:-1] = f[1:-1, :-2] + f[1:-1, 2:] + f[:-2, 1:-1] + f[2:, 1:-1]
    r = nsum - (4.0 + lam) * f - (rhs - lam * anchor.astype(np.float64))
    return float(np.sqrt(np.sum(r[cells.astype(bool)] ** 2)))
