"""Learned warm start: training pairs from the plain solver, GPU training of
the U-Net by the library's own forward/backward kernels, and the weights the
solver's ``PH_FLAG_WARM_START_NET`` path runs with.

The warm start has no counterpart in the reference (paintop's solver starts
from the anchor, crates/paintop-cpu/src/poisson.rs:169-170); the solve that
follows it is red-black SOR with the reference's relative residual
||r|| / ||r(anchor start)|| (poisson.rs:172) and stop rules.

What the network predicts: with P = the source image, the blend's solution is
u = P + h where h is harmonic inside the mask with boundary values
anchor - P on the Dirichlet ring. The network maps the f x f blocks of
(rhs, boundary values on the Dirichlet ring, interior fraction, ring density)
on a 128 x 128 grid around the mask to the block means of the solution over
interior and ring cells; the solver starts from x = P + bilinear(h). The same
network then serves as the coarse level of the solve: after each cycle of
sweeps it solves the error equation Laplacian(e) = g (rhs = block sums of the
residual, no boundary values) and the solver adds a Galerkin-scaled
bilinear(e). Training mixes both kinds of problem (``--correction-frac``);
every example is normalised by the RMS of its rhs and boundary channels, as
the solver normalises its inputs.

Usage::

    python -m paintop_amd.warmstart train --samples 60000 --epochs 30 \
        --correction-frac 0.5 --out paintop_amd/weights
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


# sweeps after which mode-2 pairs stop the learned solver (one value per chunk)
STATE_SWEEPS = (16, 24, 32, 48, 64, 96, 128, 192)
# mode-2 states closer to the solution than this relative residual are dropped:
# e = u - x is then too small against the f32 rounding of u and x
STATE_MIN_REL = 1e-4


def generate_pairs(n: int, first: int = 0, chunk: int = 96, tol: float = 1e-7, log=print,
                   mode: int = 0):
    """Solve instances first .. first+n-1 with the plain solver and return the
    coarse (inputs [n,R,R,4], target [n,R,R], weight [n,R,R]) as host arrays,
    plus the solver's sweep counts.

    mode 0: warm-start pairs (fbm source and target).
    mode 1: coarse-correction pairs -- the blend of an fbm source into a zero
    target solves Laplacian(u) = Laplacian(source), u = 0 on the ring, which is
    the correction problem Laplacian(e) = g, e = 0 on the ring, with g = the
    source's Laplacian.
    mode 2: correction pairs from solver states -- the learned solver (warm
    start and corrections, with the weights currently loaded) is stopped after
    m sweeps (x); e = u - x against the converged u is the error a correction
    at that point has to remove (e = 0 on the ring, Laplacian(e) = g there).
    These are coarsened as a mode-1 problem with source e."""
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
            if mode != 1:
                hip.fbm_device(tgt.ptr, s, s, sp["freq"], sp["seed_t"], 6, 2.0, 0.5, scale=0.5,
                               bias=0.5)
            else:
                tgt.zero()
            kind, geo = sp["shape"]
            if kind == "polygon":
                hip.polygon_mask_device(msk.ptr, s, s, geo)
            else:
                hip.ellipse_mask_device(msk.ptr, s, s, *geo)
            bufs.append((src, tgt, msk, out))
            items.append((s, s, 1, src.ptr, tgt.ptr, msk.ptr, out.ptr))
        reps = hip.poisson_solve_batch_device(desc, items, want_history=False)
        part_ok = [True] * len(specs)
        if mode == 2:
            # the learned solver's state after m sweeps, then e = u - x in the
            # source buffer, a zero anchor, and e as the solution
            m = STATE_SWEEPS[(first + c0) // chunk % len(STATE_SWEEPS)]
            xs = [DeviceArray((sp["size"], sp["size"]), np.float32) for sp in specs]
            d2 = hip.make_desc(1, 1, 1, m, tol, flags=hip.PH_FLAG_WARM_START_NET)
            items2 = [(w_, h_, c_, s_, t_, m_, x.ptr)
                      for (w_, h_, c_, s_, t_, m_, _o), x in zip(items, xs)]
            reps2 = hip.poisson_solve_batch_device(d2, items2, want_history=False)
            for j, (sp, (src, tgt, msk, out), x, r2) in enumerate(zip(specs, bufs, xs, reps2)):
                e = out.numpy().astype(np.float64) - x.numpy().astype(np.float64)
                src.copy_from(DeviceArray.from_numpy(e.astype(np.float32)))
                tgt.zero()
                out.copy_from(src)
                part_ok[j] = (not r2.converged) and r2.final_residual >= STATE_MIN_REL
                x.free()
        din = DeviceArray((len(specs), RES, RES, 4), np.float32)
        dt = DeviceArray((len(specs), RES, RES), np.float32)
        dw = DeviceArray((len(specs), RES, RES), np.float32)
        for j, (sp, (src, tgt, msk, out)) in enumerate(zip(specs, bufs)):
            s = sp["size"]
            hip.warm_coarsen_device(src.ptr, tgt.ptr, msk.ptr, out.ptr, s, s, 1, 0, 0.0,
                                    sp["x0"], sp["y0"], sp["f"], RES, din.ptr + j * R2 * 16,
                                    dt.ptr + j * R2 * 4, dw.ptr + j * R2 * 4,
                                    mode=min(mode, 1))
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
            if not part_ok[j]:
                Wt[c0 + j] = 0.0  # a state too close to the solution (mode 2)
            elif not (r.converged or r.final_residual <= 10 * tol):
                Wt[c0 + j] = 0.0
                dropped.append((first + c0 + j, r.stop_reason, r.final_residual))
                log(f"[warmstart] instance {first + c0 + j} dropped: {r.stop_reason} at "
                    f"{r.final_residual:.2e} after {r.iterations} sweeps")
        log(f"[warmstart] pairs {min(n, c0 + chunk)}/{n}  {time.time() - t0:.1f}s  "
            f"sweeps median {int(np.median(sweeps[:c0 + len(specs)]))}")
    return X, Y, Wt, sweeps


def normalise(X: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """Scale each example as the solver scales its net inputs (the problems are
    linear): channels 0 and 1 and the target are divided by the RMS of
    channels 0 and 1 over the grid (1 when both are zero). In place; returns
    the scales."""
    ms = np.einsum("nijc,nijc->n", X[..., :2].astype(np.float64),
                   X[..., :2].astype(np.float64)) / (RES * RES)
    sc = np.sqrt(ms).astype(np.float32)
    sc[sc == 0] = 1.0
    X[..., :2] /= sc[:, None, None, None]
    Y /= sc[:, None, None]
    return sc


def loss_weight(X: np.ndarray, Wt: np.ndarray) -> np.ndarray:
    """Per-cell loss weight: the coarsening's weight (1 on blocks with interior
    cells, 0.5 on blocks holding only Dirichlet-ring cells, 0 elsewhere)."""
    return np.ascontiguousarray(Wt, dtype=np.float32)


# ---------------------------------------------------------------------------
# training

def train(X, Y, Wl, *, epochs=24, batch=32, lr=2e-3, lr_min=5e-5, warmup=300, seed=1,
          val_frac=0.05, ckpt=None, log=print, time_budget=None, pyramid=1.0, augment=True,
          init=None):
    """AdamW on the GPU with a cosine schedule; each training example is drawn
    under a random one of the 8 dihedral maps of the grid (augment). Returns
    (params, history)."""
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
    def gather(idx, modes=None):
        hip.unet_gather_device(dX.ptr, dY.ptr, dW.ptr, n, idx, modes, RES, bX.ptr, bY.ptr,
                               bW.ptr)

    net = hip.UNet(batch, RES, params=init, seed=seed)
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

    def val_loss(idx=None):
        idx = np.arange(ntr, n) if idx is None else idx
        tot, cnt = 0.0, 0
        for v0 in range(0, len(idx) - batch + 1, batch):
            gather(idx[v0:v0 + batch])
            tot += net.train_step(bX.ptr, bY.ptr, bW.ptr, batch, 0.0, update=False,
                                  pyramid=pyramid)
            cnt += 1
        return tot / max(cnt, 1)

    # validation split by kind (a correction problem has no boundary values),
    # each against the loss of a zero prediction
    vidx = np.arange(ntr, n)
    is_corr = ~np.any(X[ntr:, :, :, 1] != 0, axis=(1, 2))
    kinds = [(nm, vidx[sel]) for nm, sel in (("warm", ~is_corr), ("corr", is_corr))
             if sel.sum() >= batch]
    zero = {nm: float(np.mean(np.sum(Wl[ix] * Y[ix] ** 2, axis=(1, 2)) / (RES * RES)))
            for nm, ix in kinds}

    def val_report():
        parts = []
        for nm, ix in kinds:
            v = val_loss(ix)
            parts.append(f"{nm} {v:.3e} ({v / max(zero[nm], 1e-30):.2e} of zero)")
        return "  ".join(parts)

    step = step0
    while step < total:
        k = step % steps_per_epoch
        if k == 0 and step != step0:
            perm = rng.permutation(ntr)
        idx = perm[k * batch:(k + 1) * batch]
        gather(idx, rng.integers(0, 8, size=batch) if augment else None)
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
            if len(kinds) > 1 and len(hist) % 4 == 1:
                log(f"[warmstart]   val by kind: {val_report()}")
            if ckpt:
                p, _, m, v, st = net.state()
                np.savez(ckpt, params=p, adam_m=m, adam_v=v, step=np.int64(st))
        if time_budget and now - t0 > time_budget:
            log(f"[warmstart] time budget reached at step {step}")
            break
    params = net.state()[0]
    final_val = val_loss()
    by_kind = val_report() if kinds else ""
    log(f"[warmstart] final val {final_val:.3e}  {by_kind}")
    net.close()
    for a in (dX, dY, dW, bX, bY, bW):
        a.free()
    return params, dict(history=hist, val_loss=final_val, val_by_kind=by_kind, steps=step)


def time_network(batch: int = 32, reps: int = 50, log=print) -> dict:
    """Wall-clock ms of the shipped network's forward pass at batches 1, 3, 8
    and `batch`, and of one training step (forward, multi-scale loss, backward,
    no update) at `batch`, on RES x RES grids of seeded random inputs."""
    from .device import DeviceArray, synchronize

    rng = np.random.default_rng(0)
    net = hip.UNet(batch, RES, params=load_weights())
    x = DeviceArray.from_numpy(rng.standard_normal((batch, RES, RES, 4)).astype(np.float32))
    y = DeviceArray.from_numpy(rng.standard_normal((batch, RES, RES)).astype(np.float32))
    w = DeviceArray.from_numpy(np.ones((batch, RES, RES), np.float32))
    o = DeviceArray((batch, RES, RES), np.float32)

    def ms(fn, n):
        for _ in range(5):
            fn()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    res = {}
    for b in sorted({1, 3, 8, batch}):
        if b <= batch:
            res[f"forward_b{b}_ms"] = round(ms(lambda: net.forward(x.ptr, o.ptr, b), reps), 4)
    res[f"train_step_b{batch}_ms"] = round(
        ms(lambda: net.train_step(x.ptr, y.ptr, w.ptr, batch, 0.0, update=False, pyramid=1.0),
           max(1, reps // 2)), 4)
    log(json.dumps(res))
    return res


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
    t.add_argument("--no-augment", action="store_true", help="no dihedral augmentation")
    t.add_argument("--correction-frac", type=float, default=0.5,
                   help="fraction of the samples that are coarse-correction problems")
    t.add_argument("--state-frac", type=float, default=0.0,
                   help="fraction of the samples that are correction problems from the "
                        "learned solver's own states (needs --init or the shipped weights)")
    t.add_argument("--init", default=None,
                   help="start from these weights (.npy); also the solver's for --state-frac")
    tm = sub.add_parser("time", help="time the network's forward pass and training step")
    tm.add_argument("--batch", type=int, default=32)
    tm.add_argument("--reps", type=int, default=50)
    a = ap.parse_args(argv)
    if a.cmd == "time":
        time_network(a.batch, a.reps)
        return 0
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.cmd == "train":
        t0 = time.time()
        n1 = int(round(a.samples * a.correction_frac))
        n2 = int(round(a.samples * a.state_frac))
        n0 = a.samples - n1 - n2
        init = load_weights(a.init) if a.init else None
        parts = [generate_pairs(n0, log=log, mode=0)] if n0 else []
        if n1:
            # correction problems on instances of their own (other masks and textures)
            parts.append(generate_pairs(n1, first=n0, log=log, mode=1))
        if n2:
            # the solver's own states, corrections on for these small crops too
            hip.warm_start_set_weights(init if init is not None else load_weights())
            os.environ["PAINTOP_HIP_CORR_MIN_CELLS"] = "0"
            parts.append(generate_pairs(n2, first=n0 + n1, log=log, mode=2))
            del os.environ["PAINTOP_HIP_CORR_MIN_CELLS"]
            hip.warm_start_set_weights(None)
        X, Y, Wt, sweeps = (np.concatenate(z) for z in zip(*parts))
        # validation examples are the last ones: interleave the two kinds
        order = np.random.default_rng(7).permutation(len(X))
        X, Y, Wt, sweeps = X[order], Y[order], Wt[order], sweeps[order]
        normalise(X, Y)
        gen_s = time.time() - t0
        Wl = loss_weight(X, Wt)
        params, info = train(X, Y, Wl, epochs=a.epochs, batch=a.batch, lr=a.lr, ckpt=a.ckpt,
                             log=log, time_budget=a.time_budget, pyramid=a.pyramid,
                             augment=not a.no_augment, init=init)
        os.makedirs(a.out, exist_ok=True)
        np.save(os.path.join(a.out, "warm_start_unet.npy"), params.astype(np.float32))
        meta = dict(samples=a.samples, correction_samples=n1, state_samples=n2,
                    init=os.path.basename(a.init) if a.init else None, epochs=a.epochs,
                    batch=a.batch,
                    lr=a.lr, input_normalisation="rms of channels 0-1",
                    pyramid=a.pyramid, augment=not a.no_augment,
                    dropped=int((Wt.reshape(len(Wt), -1).max(axis=1) == 0).sum()),
                    steps=info["steps"], val_loss=info["val_loss"],
                    val_by_kind=info["val_by_kind"],
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
