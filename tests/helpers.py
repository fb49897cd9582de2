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


def star_case(n):
    """The 4096^2 seamless-clone workload scaled to n^2 (star mask, fbm with the
    same features relative to the image): cells, rhs, anchor."""
    import math

    from oracle import oracle
    from paintop_amd.inputs import star_polygon

    sc = n / 4096
    verts = [(x * sc, y * sc) for (x, y) in
             star_polygon(2048, 2048, math.sqrt(0.3 * 4096 * 4096 / 1.5), 0.5, 6, 0.1, 301)]
    cells = classify(oracle.polygon_mask(n, n, verts))
    s = oracle.fbm(n, n, 0.125 / sc, 101, 6) * np.float32(0.5) + np.float32(0.5)
    t = oracle.fbm(n, n, 0.125 / sc, 201, 6) * np.float32(0.5) + np.float32(0.5)
    return cells, np.where(cells == 1, laplacian(s), 0.0), t.astype(np.float64)


def smallest_eigenvalue(cells):
    """Smallest eigenvalue of the 5-point operator 4I - N on the interior cells
    (scipy shift-invert Lanczos)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    ys, xs = np.nonzero(cells)
    idx = -np.ones(cells.shape, np.int64)
    idx[ys, xs] = np.arange(ys.size)
    rows, cols = [np.arange(ys.size)], [np.arange(ys.size)]
    vals = [np.full(ys.size, 4.0)]
    for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        nb = idx[ys + dy, xs + dx]
        k = nb >= 0
        rows.append(np.arange(ys.size)[k])
        cols.append(nb[k])
        vals.append(-np.ones(int(k.sum())))
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))))
    return float(spla.eigsh(A, k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0])
