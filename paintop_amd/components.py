"""A Poisson blend solved per connected component of its mask, all components
in one batched call.

The reference solves the whole mask as one system (crates/paintop-cpu/src/
poisson_blend.rs:168-236) over the bounding box of every interior cell; it
notes solving per connected mask component as the way to shrink that work
(RESEARCH.md:228), and an executor that gathers many masked regions into one
persistent launch is the batched entry point's purpose
(``ph_poisson_solve_batch_device``).

Why the split solves the same system. Interior cells are those with
mask > 0.5 off the image border (poisson.rs:340-355). Two interior cells of
different 4-connected components are never 4-neighbours, so each component's
equations involve only its own unknowns and fixed (non-interior) cells: the
blend is a set of independent systems. Cropping a component's bounding box
plus its one-cell Dirichlet ring, with the crop's mask set to that component
alone, gives exactly its system (other components' cells inside the crop are
never neighbours of its interior cells).

Stop rule. Each component stops at its own relative residual
||r_k|| <= tol * ||r_k(anchor start)||; then the whole mask's residual
satisfies ||r||^2 = sum ||r_k||^2 <= tol^2 * sum r0_k^2 = tol^2 * r0^2, the
reference's criterion (poisson.rs:172, 202-209) for the whole mask. The
iteration counts and residual histories are the components' own: the report
is the worst component's (more iterations, then the larger final residual,
as the worst channel is chosen, poisson_blend.rs:240-246), with every
component's report beside it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import hip

MASK_THRESHOLD = 0.5  # poisson.rs:329


def interior_cells(mask: np.ndarray) -> np.ndarray:
    """classify_cells (poisson.rs:340-355): mask > 0.5 and not on the image
    border; every cell is a boundary cell when width or height is below 3."""
    h, w = mask.shape
    cls = np.zeros((h, w), dtype=bool)
    if h >= 3 and w >= 3:
        cls[1:-1, 1:-1] = mask[1:-1, 1:-1] > MASK_THRESHOLD
    return cls


@dataclass
class Component:
    label: int
    y0: int
    y1: int  # exclusive
    x0: int
    x1: int  # exclusive
    cells: int


def components(mask: np.ndarray) -> tuple[np.ndarray, list[Component]]:
    """4-connected components of the interior cells, in row-major order of
    their first cell. Returns (labels, components): labels is 0 outside the
    interior and 1..n inside; each component's box includes its Dirichlet ring."""
    from scipy import ndimage

    cls = interior_cells(mask)
    cross = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=bool)
    labels, n = ndimage.label(cls, structure=cross)
    out = []
    for k, sl in enumerate(ndimage.find_objects(labels), start=1):
        ys, xs = sl
        out.append(Component(k, ys.start - 1, ys.stop + 1, xs.start - 1, xs.stop + 1,
                             int((labels[sl] == k).sum())))
    return labels.astype(np.int32), out


def solve_components(desc: hip.PhSolveDesc, source: np.ndarray, target: np.ndarray,
                     mask: np.ndarray, stream: int = 0):
    """Solve the blend of (source, target, mask) per connected component in one
    batched call. source/target are (H, W, C) or (H, W) f32 host arrays, mask
    (H, W). Returns (candidate, worst report, [(component, report)]); the
    candidate is the target with every component's interior replaced by its
    solution (outside-mask samples keep the target's bits)."""
    from .device import DeviceArray

    src = np.ascontiguousarray(source, dtype=np.float32)
    tgt = np.ascontiguousarray(target, dtype=np.float32)
    msk = np.ascontiguousarray(mask, dtype=np.float32)
    squeeze = src.ndim == 2
    if squeeze:
        src, tgt = src[..., None], tgt[..., None]
    h, w, c = src.shape
    labels, comps = components(msk)
    cand = tgt.copy()
    if not comps:
        return (cand[..., 0] if squeeze else cand), None, []
    bufs, items = [], []
    for cp in comps:
        ys, xs = slice(cp.y0, cp.y1), slice(cp.x0, cp.x1)
        m = (labels[ys, xs] == cp.label).astype(np.float32)
        arrs = [DeviceArray.from_numpy(np.ascontiguousarray(src[ys, xs])),
                DeviceArray.from_numpy(np.ascontiguousarray(tgt[ys, xs])),
                DeviceArray.from_numpy(m),
                DeviceArray((cp.y1 - cp.y0, cp.x1 - cp.x0, c), np.float32)]
        bufs.append(arrs)
        items.append((cp.x1 - cp.x0, cp.y1 - cp.y0, c, arrs[0].ptr, arrs[1].ptr, arrs[2].ptr,
                      arrs[3].ptr))
    try:
        reps = hip.poisson_solve_batch_device(desc, items, stream)
        for cp, arrs in zip(comps, bufs):
            ys, xs = slice(cp.y0, cp.y1), slice(cp.x0, cp.x1)
            sol = arrs[3].numpy()
            sel = labels[ys, xs] == cp.label
            cand[ys, xs][sel] = sol[sel]
    finally:
        for arrs in bufs:
            for a in arrs:
                a.free()
    worst = max(reps, key=lambda r: (r.iterations, r.final_residual))
    return (cand[..., 0] if squeeze else cand), worst, list(zip(comps, reps))
