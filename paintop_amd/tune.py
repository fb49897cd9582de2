"""Sweep-configuration tuner: half-sweeps per launch (K) and tile rows (hb).

Times plain and learned solves of the seamless-clone generator (RGBA star
masks, 256^2 .. 4096^2, tolerance 1e-6) under every candidate configuration on
the device, picks the fastest configuration per problem, and writes
``paintop_amd/csrc/sor_tuning.h``: for each form, rows of (work bound, K, hb),
where a call's work is the sum over its images of channels x the area of the
interior's bounding box (the solver computes the same figure after its
classification and takes the first row whose bound is not below it). The
fields a configuration leaves are bit-identical (the sweep kernel's schedule
reads exactly what a global red/black sweep reads), so the table changes
speed only. The timing table goes to stdout (markdown) and, with --report, to
a file.

Usage (on the device): ``python -m paintop_amd.tune [--out FILE] [--report FILE]``
then rebuild the library (``python -m paintop_amd.build``).
"""
from __future__ import annotations

import argparse
import math
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "csrc", "sor_tuning.h")

# (K, tile rows): K = 16 with tiles of two ring periods for large calls, down to
# K = 8 with short tiles for small ones; the tile rows are even and >= 16
CONFIGS = [(16, 40), (16, 20), (12, 32), (12, 16), (8, 24), (8, 16)]
SIZES = [256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096]
# configurations within this fraction of a problem's fastest count as equal
SLACK = 0.05


def interior_bbox_area(mask: np.ndarray) -> int:
    """Area of the bounding box of the interior cells (mask > 0.5, not on the
    image border; poisson.rs:329-355), as the solver's classification finds it."""
    h, w = mask.shape
    inside = np.zeros_like(mask, dtype=bool)
    inside[1:h - 1, 1:w - 1] = mask[1:h - 1, 1:w - 1] > 0.5
    ys, xs = np.nonzero(inside)
    if len(ys) == 0:
        return 0
    return int((ys.max() - ys.min() + 1) * (xs.max() - xs.min() + 1))


def time_solve(size: int, k: int, hb: int, learned: bool, reps: int) -> tuple[float, int]:
    """Median wall ms of `reps` solves (after one untimed) and the sweep count."""
    from paintop_amd import hip
    from paintop_amd.device import DeviceArray, synchronize
    from paintop_amd.inputs import seamless_clone_case

    src, tgt, msk = seamless_clone_case(size, channels=4)
    out = DeviceArray(tgt.shape, np.float32)
    flags = hip.PH_FLAG_WARM_START_NET if learned else 0
    desc = hip.make_desc(size, size, 4, 100000, 1e-6, 0.0, hip.PH_OP_POISSON_BLEND,
                         half_sweeps=k, tile_rows=hb, flags=flags)
    try:
        hip.poisson_solve_device(desc, src.ptr, tgt.ptr, msk.ptr, out.ptr, want_history=False)
        times, sweeps = [], 0
        for _ in range(reps):
            synchronize()
            t0 = time.perf_counter()
            r = hip.poisson_solve_device(desc, src.ptr, tgt.ptr, msk.ptr, out.ptr,
                                         want_history=False)
            synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
            sweeps = r.iterations
        return float(np.median(times)), sweeps
    finally:
        for a in (src, tgt, msk, out):
            a.free()


def table_rows(works: list[float], times: list[list[float]],
               slack: float = SLACK) -> list[tuple[float, int, int]]:
    """Rows (work bound, K, hb) from the problems' timings or scores (one list
    per problem, in CONFIGS order; lower is better): a configuration is
    acceptable for a problem when within `slack` of its best; going up in work, each row takes the
    acceptable configuration that stays acceptable for the most problems (ties:
    the faster one here), so timing noise between near-equal configurations
    does not split the table. A row ends at the geometric mean of its last
    problem's work and the next one's."""
    order = [int(i) for i in np.argsort(works)]
    w = [works[i] for i in order]
    t = [times[i] for i in order]
    ok = [[x <= (1.0 + slack) * min(ts) for x in ts] for ts in t]
    rows = []
    i = 0
    while i < len(w):
        def run(c):
            j = i
            while j < len(w) and ok[j][c]:
                j += 1
            return j - i
        cands = [c for c in range(len(CONFIGS)) if ok[i][c]]
        c = max(cands, key=lambda c: (run(c), -t[i][c]))
        j = i + run(c) - 1
        bound = math.sqrt(w[j] * w[j + 1]) if j + 1 < len(w) else 1e300
        rows.append((bound, CONFIGS[c][0], CONFIGS[c][1]))
        i = j + 1
    return rows


def score(plain_ms: list[float], learned_ms: list[float]) -> list[float]:
    """Per configuration: the geometric mean of its plain and learned times,
    each relative to that form's fastest configuration."""
    bp, bl = min(plain_ms), min(learned_ms)
    return [math.sqrt((p / bp) * (q / bl)) for p, q in zip(plain_ms, learned_ms)]


def write_header(path: str, rows, device: str) -> None:
    def fmt(rows):
        return "\n".join(f"    {{{bound:.3e}, {k}, {hb}}}," for bound, k, hb in rows)

    text = f"""// sor_tuning.h -- half-sweeps per launch (K) and sweep-tile rows (hb) by the
// size of a call, written by the tuner (python -m paintop_amd.tune) from timed
// plain and learned solves of the seamless-clone generator
// (device: {device}). Work W of a call = the sum over its images of
// channels x the area of the interior's bounding box; the first row with
// W <= w_max applies. Plain and learned calls share the table.
#pragma once

struct PhSorTune {{
  double w_max;
  int k, hb;
}};

static const PhSorTune PH_SOR_TUNING[] = {{
{fmt(rows)}
}};
"""
    with open(path, "w") as f:
        f.write(text)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=HEADER, help="header to write")
    ap.add_argument("--report", default=None, help="markdown timing table to write")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--sizes", type=int, nargs="*", default=SIZES)
    args = ap.parse_args()

    from paintop_amd import hip, warmstart
    from paintop_amd.inputs import seamless_clone_case

    info = hip.probe()
    warmstart.enable()
    lines = [f"# Sweep configuration timings ({info['name']}, seamless-clone RGBA star, "
             f"tol 1e-6, median of {args.reps})", "",
             "| size | work W | form | " + " | ".join(f"K={k} hb={hb}" for k, hb in CONFIGS) +
             " | best |", "|---" * (len(CONFIGS) + 4) + "|"]
    works, timings = {False: [], True: []}, {False: [], True: []}
    for n in args.sizes:
        _, _, msk = seamless_clone_case(n, channels=4)
        w = 4.0 * interior_bbox_area(msk.numpy())
        msk.free()
        for learned in (False, True):
            ms = []
            for k, hb in CONFIGS:
                t, sweeps = time_solve(n, k, hb, learned, args.reps)
                ms.append(t)
            i = int(np.argmin(ms))
            works[learned].append(w)
            timings[learned].append(ms)
            cells = " | ".join(f"**{t:.2f}**" if j == i else f"{t:.2f}" for j, t in enumerate(ms))
            lines.append(f"| {n} | {w:.3e} | {'learned' if learned else 'plain'} | {cells} | "
                         f"K={CONFIGS[i][0]} hb={CONFIGS[i][1]} |")
            print(lines[-1], flush=True)
    # one table for both forms: a call asking for the learned start whose start
    # loses (or is dropped) must give the plain call's result bit for bit, and
    # the launch boundaries (refinement points, the race's probe) depend on K.
    # A configuration's score is the geometric mean of its slowdowns against
    # each form's fastest.
    scores = [score(tp, tl) for tp, tl in zip(timings[False], timings[True])]
    rows = table_rows(works[False], scores)
    write_header(args.out, rows, info["name"])
    lines += ["", "score (geometric mean of the slowdowns against each form's fastest):"]
    for n, w, sc in zip(args.sizes, works[False], scores):
        lines.append(f"  {n}: " + " ".join(f"{x:.3f}" for x in sc))
    lines += ["", f"rows: {rows}", f"written: {os.path.relpath(args.out, os.path.dirname(HERE))}"]
    text = "\n".join(lines) + "\n"
    print(text)
    if args.report:
        with open(args.report, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
