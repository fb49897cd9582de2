"""Learned warm start: training pairs from the plain solver, GPU training of
the U-Net by the library's own forward/backward kernels, and the weights the
solver's ``PH_FLAG_WARM_START_NET`` path runs with.

The warm start has no counterpart in the reference (paintop's solver starts
from the anchor, crates/paintop-cpu/src/poisson.rs:169-170); the solve that
follows it is the unchanged red-black SOR with the reference's relative
residual ||r|| / ||r(anchor start)|| (poisson.rs:172) and stop rules.

What the network predicts: with P = the source image, the blend's solution is
u = P + h where h is harmonic inside the mask with boundary values
anchor - P on the Dirichlet ring. The network maps the f x f blocks of
(rhs of the correction problem, boundary values on the Dirichlet ring,
interior fraction, ring density) on a 128 x 128 grid around the mask to the
block means of h over interior and ring cells; the solver starts from
x = P + bilinear(h).

Usage::

    python -m paintop_amd.warmstart train --samples 20000 --epochs 24 \\
        --out paintop_amd/weights
    from paintop_amd import warmstart; warmstart.enable()   # load the shipped weights
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import hip
from .inputs import SplitMix64, random_polygon, star_polygon

RES = 128  # the solver's coarse grid (paintop_amd/csrc/unet.hip WARM_RES)
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(HERE, "weights")
WEIGHTS_FILE = os.path.join(WEIGHTS_DIR, "warm_start_unet.npy")
SEED0 = 10_000  # training instance i draws from SplitMix64(SEED0 + i)


# ---------------------------------------------------------------------------
# weights

def load_weights(path: str = WEIGHTS_FILE) -> np.ndarray:
    a = np.load(path, allow_pickle=False)
    if a.dtype != np.float32 or a.ndim != 1 or a.size != hip.unet_param_storage():
        raise ValueError(f"{path}: expected {hip.unet_param_storage()} float32 values, "
                         f"got {a.dtype} {a.shape}")
    return a


def enable(path: str = WEIGHTS_FILE) -> np.ndarray:
    """Load the weights and hand them to the solver's warm start."""
    w = load_weights(path)
    hip.warm_start_set_weights(w)
    return w


def disable() -> None:
    hip.warm_start_set_weights(None)


# ---------------------------------------------------------------------------
# training pairs

def _region(bx0: int, by0: int, bx1: int, by1: int):
    """The solver's coarse-grid framing of an interior bounding box (inclusive
    cell bounds): the plane region (box + Dirichlet ring, even-parity origin)
    centred in a RES*f square (solver.hip, warm start)."""
    rx0, ry0 = bx0 - 1, by0 - 1
    if (rx0 + ry0) & 1:
        rx0 -= 1
    rw, rh = bx1 + 2 - rx0, by1 + 2 - ry0
    f = max(1, (max(rw, rh) + RES - 1) // RES)
    return rx0 + rw // 2 - (RES * f) // 2, ry0 + rh // 2 - (RES * f) // 2, f


def _polygon_area(verts):
    x = np.array([v[0] for v in verts])
    y = np.array([v[1] for v in verts])
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def sample_spec(i: int) -> dict:
    """Instance i of the training set: a square crop of 256..1024 px with a star
    (12 vertices), polygon (6-10 vertices) or ellipse mask filling 15-55 % of
    it, and fbm source/target (6 octaves, gain 0.5) whose base wavelength spans
    6..64 cells of the coarse grid."""
    rng = SplitMix64(SEED0 + i)
    s = 128 * (2 + int(rng.uniform() * 7) % 7)
    fill = rng.uniform(0.15, 0.55)
    kind = rng.uniform()
    cx = s / 2.0 + rng.uniform(-0.08, 0.08) * s
    cy = s / 2.0 + rng.uniform(-0.08, 0.08) * s
    if kind < 0.35:
        inner = rng.uniform(0.35, 0.65)
        verts = star_polygon(0.0, 0.0, 1.0, inner, 6, 0.1, SEED0 + 7919 * i)
        rot = rng.uniform(0.0, 2.0 * math.pi)
        c, sn = math.cos(rot), math.sin(rot)
        verts = [(x * c - y * sn, x * sn + y * c) for x, y in verts]
        shape = ("polygon", verts)
    elif kind < 0.7:
        nv = 6 + int(rng.uniform() * 5) % 5
        ar = rng.uniform(0.6, 1.6)
        verts = random_polygon(0.0, 0.0, math.sqrt(ar), 1.0 / math.sqrt(ar), nv, rng)
        shape = ("polygon", verts)
    else:
        ar = rng.uniform(0.5, 2.0)
        shape = ("ellipse", (math.sqrt(ar), 1.0 / math.sqrt(ar)))
    # scale to the fill fraction, then shrink if it would leave the crop
    if shape[0] == "polygon":
        sc = math.sqrt(fill * s * s / max(_polygon_area(shape[1]), 1e-9))
        ext = max(max(abs(x), abs(y)) for x, y in shape[1]) * sc
    else:
        rx, ry = shape[1]
        sc = math.sqrt(fill * s * s / (math.pi * rx * ry))
        ext = max(rx, ry) * sc
    lim = min(cx, cy, s - cx, s - cy) - 3.0
    if ext > lim:
        sc *= lim / ext
    if shape[0] == "polygon":
        verts = [(cx + sc * x, cy + sc * y) for x, y in shape[1]]
        xs, ys = [v[0] for v in verts], [v[1] for v in verts]
        box = (min(xs), min(ys), max(xs), max(ys))
        shape = ("polygon", verts)
    else:
        rx, ry = shape[1][0] * sc, shape[1][1] * sc
        box = (cx - rx, cy - ry, cx + rx, cy + ry)
        shape = ("ellipse", (cx, cy, rx, ry))
    # interior cells have centres inside: cell x with x + 0.5 in [lo, hi]
    bx0 = max(1, int(math.ceil(box[0] - 0.5)))
    by0 = max(1, int(math.ceil(box[1] - 0.5)))
    bx1 = min(s - 2, int(math.floor(box[2] - 0.5)))
    by1 = min(s - 2, int(math.floor(box[3] - 0.5)))
    x0, y0, f = _region(bx0, by0, bx1, by1)
    wl_coarse = math.exp(rng.uniform(math.log(6.0), math.log(64.0)))
    freq = 64.0 / (wl_coarse * f)  # fbm scale = frequency / 64 per pixel
    return dict(size=s, shape=shape, x0=x0, y0=y0, f=f, freq=freq,
                seed_s=2 * (SEED0 + i) + 1_000_001, seed_t=2 * (SEED0 + i) + 1_000_002)


def generate_pairs(n: int, first: int = 0, chunk: int = 96, tol: float = 1e-7, log=print):
    """Solve instances first .. first+n-1 with the plain solver and return the
    coarse (inputs [n,R,R,4], target [n,R,R], weight [n,R,R]) as host arrays,
    plus the solver's sweep counts."""
    from .device import DeviceArray, synchronize

    R2 = RES * RES
    X = np.zeros((n, RES, RES, 4), np.float32)
    Y = np.zeros((n, RES, RES), np.float32)
    Wt = np.zeros((n, RES, RES), np.float32)
    sweeps = np.zeros(n, np.int64)
    dropped = []
    desc = hip.make_desc(1, 1, 1, 100000, tol)
    t0 = time.time()
    for c0 in range(0, n, chunk):
        specs = [sample_spec(first + j) for j in range(c0, min(n, c0 + chunk))]
        bufs, items = [], []
        for sp in specs:
            s = sp["size"]
            src, tgt, msk, out = (DeviceArray((s, s), np.float32) for _ in range(4))
            hip.fbm_device(src.ptr, s, s, sp["freq"], sp["seed_s"], 6, 2.0, 0.5, scale=0.5,
                           bias=0.5)
            hip.fbm_device(tgt.ptr, s, s, sp["freq"], sp["seed_t"], 6, 2.0, 0.5, scale=0.5,
                           bias=0.5)
            kind, geo = sp["shape"]
            if kind == "polygon":
                hip.polygon_mask_device(msk.ptr, s, s, geo)
            else:
                hip.ellipse_mask_device(msk.ptr, s, s, *geo)
            bufs.append((src, tgt, msk, out))
            items.append((s, s, 1, src.ptr, tgt.ptr, msk.ptr, out.ptr))
        reps = hip.poisson_solve_batch_device(desc, items, want_history=False)
        din = DeviceArray((len(specs), RES, RES, 4), np.float32)
        dt = DeviceArray((len(specs), RES, RES), np.float32)
        dw = DeviceArray((len(specs), RES, RES), np.float32)
        for j, (sp, (src, tgt, msk, out)) in enumerate(zip(specs, bufs)):
            s = sp["size"]
            hip.warm_coarsen_device(src.ptr, tgt.ptr, msk.ptr, out.ptr, s, s, 1, 0, 0.0,
                                    sp["x0"], sp["y0"], sp["f"], RES, din.ptr + j * R2 * 16,
                                    dt.ptr + j * R2 * 4, dw.ptr + j * R2 * 4)
        synchronize()
        k = slice(c0, c0 + len(specs))
        X[k], Y[k], Wt[k] = din.numpy(), dt.numpy(), dw.numpy()
        sweeps[k] = [r.iterations for r in reps]
        for b in bufs:
            for a in b:
                a.free()
        for a in (din, dt, dw):
            a.free()
        for j, r in enumerate(reps):
            # a solve that ends (stalled) within 10x the tolerance is still a
            # clean target; anything else is dropped (zero loss weight)
            if not (r.converged or r.final_residual <= 10 * tol):
                Wt[c0 + j] = 0.0
                dropped.append((first + c0 + j, r.stop_reason, r.final_residual))
                log(f"[warmstart] instance {first + c0 + j} dropped: {r.stop_reason} at "
                    f"{r.final_residual:.2e} after {r.iterations} sweeps")
        log(f"[warmstart] pairs {min(n, c0 + chunk)}/{n}  {time.time() - t0:.1f}s  "
            f"sweeps median {int(np.median(sweeps[:c0 + len(specs)]))}")
    return X, Y, Wt, sweeps


def loss_weight(X: np.ndarray, Wt: np.ndarray) -> np.ndarray:
    """Per-cell loss weight: the coarsening's weight (1 on blocks with interior
    cells, 0.5 on blocks holding only Dirichlet-ring cells, 0 elsewhere)."""
    return np.ascontiguousarray(Wt, dtype=np.float32)


# ---------------------------------------------------------------------------
# training

def train(X, Y, Wl, *, epochs=24, batch=32, lr=2e-3, lr_min=5e-5, warmup=300, seed=1,
          val_frac=0.05, ckpt=None, log=print, time_budget=None, pyramid=1.0):
    """AdamW on the GPU with a cosine schedule. Returns (params, history)."""
    from .device import DeviceArray

    n = X.shape[0]
    nval = max(1, int(n * val_frac))
    ntr = n - nval
    R2 = RES * RES
    dX = DeviceArray.from_numpy(X)
    dY = DeviceArray.from_numpy(Y)
    dW = DeviceArray.from_numpy(Wl)
    bX = DeviceArray((batch, RES, RES, 4), np.float32)
    bY = DeviceArray((batch, RES, RES), np.float32)
    bW = DeviceArray((batch, RES, RES), np.float32)
    from .device import runtime, D2D

    rt = runtime()

    def gather(idx):
        for j, i in enumerate(idx):
            rt.hipMemcpy(bX.ptr + j * R2 * 16, dX.ptr + int(i) * R2 * 16, R2 * 16, D2D)
            rt.hipMemcpy(bY.ptr + j * R2 * 4, dY.ptr + int(i) * R2 * 4, R2 * 4, D2D)
            rt.hipMemcpy(bW.ptr + j * R2 * 4, dW.ptr + int(i) * R2 * 4, R2 * 4, D2D)

    net = hip.UNet(batch, RES, seed=seed)
    step0 = 0
    rng = np.random.default_rng(seed)
    if ckpt and os.path.exists(ckpt):
        z = np.load(ckpt, allow_pickle=False)
        net.set_state(z["params"], z["adam_m"], z["adam_v"], int(z["step"]))
        step0 = int(z["step"])
        rng = np.random.default_rng(seed + step0)
        log(f"[warmstart] resumed at step {step0}")
    steps_per_epoch = ntr // batch
    total = epochs * steps_per_epoch
    hist = []
    t0 = time.time()
    perm = rng.permutation(ntr)
    last = t0

    def val_loss():
        tot, cnt = 0.0, 0
        for v0 in range(ntr, n - batch + 1, batch):
            gather(range(v0, v0 + batch))
            tot += net.train_step(bX.ptr, bY.ptr, bW.ptr, batch, 0.0, update=False,
                                  pyramid=pyramid)
            cnt += 1
        return tot / max(cnt, 1)

    step = step0
    while step < total:
        k = step % steps_per_epoch
        if k == 0 and step != step0:
            perm = rng.permutation(ntr)
        idx = perm[k * batch:(k + 1) * batch]
        gather(idx)
        a = min(1.0, (step + 1) / warmup)
        cosf = 0.5 * (1 + math.cos(math.pi * step / max(total, 1)))
        cur = lr_min + (lr - lr_min) * cosf
        loss = net.train_step(bX.ptr, bY.ptr, bW.ptr, batch, cur * a, weight_decay=1e-4,
                              pyramid=pyramid)
        step += 1
        now = time.time()
        if now - last > 30 or step == total:
            last = now
            vl = val_loss()
            hist.append((step, loss, vl))
            log(f"[warmstart] step {step}/{total} lr {cur * a:.2e} loss {loss:.3e} val {vl:.3e} "
                f"{now - t0:.0f}s")
            if ckpt:
                p, _, m, v, st = net.state()
                np.savez(ckpt, params=p, adam_m=m, adam_v=v, step=np.int64(st))
        if time_budget and now - t0 > time_budget:
            log(f"[warmstart] time budget reached at step {step}")
            break
    params = net.state()[0]
    final_val = val_loss()
    net.close()
    for a in (dX, dY, dW, bX, bY, bW):
        a.free()
    return params, dict(history=hist, val_loss=final_val, steps=step)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m paintop_amd.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="generate solver pairs and train on the GPU")
    t.add_argument("--samples", type=int, default=20000)
    t.add_argument("--epochs", type=int, default=24)
    t.add_argument("--batch", type=int, default=32)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", default=WEIGHTS_DIR)
    t.add_argument("--ckpt", default=None, help="checkpoint file (resume if present)")
    t.add_argument("--time-budget", type=float, default=None, help="seconds of training")
    t.add_argument("--pyramid", type=float, default=1.0, help="multi-scale loss weight")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.cmd == "train":
        t0 = time.time()
        X, Y, Wt, sweeps = generate_pairs(a.samples, log=log)
        gen_s = time.time() - t0
        Wl = loss_weight(X, Wt)
        params, info = train(X, Y, Wl, epochs=a.epochs, batch=a.batch, lr=a.lr, ckpt=a.ckpt,
                             log=log, time_budget=a.time_budget, pyramid=a.pyramid)
        os.makedirs(a.out, exist_ok=True)
        np.save(os.path.join(a.out, "warm_start_unet.npy"), params.astype(np.float32))
        meta = dict(samples=a.samples, epochs=a.epochs, batch=a.batch, lr=a.lr,
                    pyramid=a.pyramid,
                    dropped=int((Wt.reshape(len(Wt), -1).max(axis=1) == 0).sum()),
                    steps=info["steps"], val_loss=info["val_loss"],
                    history=info["history"][-20:], pair_generation_s=round(gen_s, 1),
                    train_s=round(time.time() - t0 - gen_s, 1),
                    solver_sweeps_median=int(np.median(sweeps)),
                    param_count=hip.unet_param_count(), res=RES, seed0=SEED0)
        with open(os.path.join(a.out, "warm_start_unet.json"), "w") as fh:
            json.dump(meta, fh, indent=1)
        log(f"[warmstart] saved {a.out}/warm_start_unet.npy  val {info['val_loss']:.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
