"""CPU reference of the warm-start U-Net: forward pass, loss and backward pass
in float64 numpy. TEST INFRASTRUCTURE ONLY: imported by tests/ to check the
HIP network kernels; the product never calls it.

The network has no counterpart in the reference repository (paintop's solver
always starts from the anchor, crates/paintop-cpu/src/poisson.rs:169-170);
this file restates the architecture documented in
paintop_amd/csrc/unet_internal.h. It is pinned by itself: the backward pass
here is written as the adjoint of each forward step (patch extraction,
nearest upsample, average pool, ReLU), independently of the GPU's
flipped-weight input-gradient convolution, and tests check it against
central finite differences of the forward pass.

With ``bf16=True`` every tensor the GPU stores in bfloat16 (weights, network
input, activations, activation gradients) is rounded to bfloat16 (round to
nearest even) at the same point, so GPU and CPU differ only by accumulation
order and precision.
"""
from __future__ import annotations

import numpy as np

# (level, in_a, in_b, ca, cb, cout, out) -- the layer table of unet_internal.h
BUFS = ["X", "A0", "E0", "P0", "A1", "E1", "P1", "A2", "E2", "P2", "A3", "B3",
        "D2A", "D2", "D1A", "D1", "D0A", "D0"]
LAYERS = [
    (0, "X", None, 4, 0, 16, "A0"),
    (0, "A0", None, 16, 0, 16, "E0"),
    (1, "P0", None, 16, 0, 32, "A1"),
    (1, "A1", None, 32, 0, 32, "E1"),
    (2, "P1", None, 32, 0, 64, "A2"),
    (2, "A2", None, 64, 0, 64, "E2"),
    (3, "P2", None, 64, 0, 64, "A3"),
    (3, "A3", None, 64, 0, 64, "B3"),
    (2, "E2", "B3", 64, 64, 64, "D2A"),
    (2, "D2A", None, 64, 0, 64, "D2"),
    (1, "E1", "D2", 32, 64, 32, "D1A"),
    (1, "D1A", None, 32, 0, 32, "D1"),
    (0, "E0", "D1", 16, 32, 16, "D0A"),
    (0, "D0A", None, 16, 0, 16, "D0"),
]
POOLS = {"E0": "P0", "E1": "P1", "E2": "P2"}
DIL = [1, 1, 1, 1, 1, 1, 2, 4, 1, 1, 1, 1, 1, 1]  # dilation of each layer's 3x3 taps
HEAD_C = 16


def kp_of(cin: int) -> int:
    return ((9 * cin + 15) // 16) * 16


def layout():
    """Offsets of the FP32 parameter storage: per layer (off_w, off_b, cout, cp,
    kp), then the head offset and the total size."""
    out, o = [], 0
    for (_, _, _, ca, cb, cout, _) in LAYERS:
        cp = ca + cb
        kp = kp_of(cp)
        out.append((o, o + cout * kp, cout, cp, kp))
        o += cout * kp + cout
    return out, o, o + HEAD_C + 1


def param_count() -> int:
    n = 0
    for (_, _, _, ca, cb, cout, _) in LAYERS:
        n += 9 * (ca + cb) * cout + cout
    return n + HEAD_C + 1


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _conv_weights(p, li, rnd):
    lay, _, _ = layout()
    ow, ob, cout, cp, kp = lay[li]
    w = p[ow:ow + cout * kp].reshape(cout, kp)[:, :9 * cp].reshape(cout, 9, cp).astype(np.float64)
    if rnd:
        w = bf16(w)
    return w, p[ob:ob + cout].astype(np.float64)


def _patches(x, dil=1):
    """x [B,H,W,C] -> [B,H,W,9,C], tap t at offset dil*(t//3 - 1, t%3 - 1),
    zero padding."""
    B, H, W, C = x.shape
    q = dil
    xp = np.zeros((B, H + 2 * q, W + 2 * q, C))
    xp[:, q:q + H, q:q + W] = x
    out = np.empty((B, H, W, 9, C))
    for t in range(9):
        dy, dx = (t // 3 - 1) * dil, (t % 3 - 1) * dil
        out[:, :, :, t] = xp[:, q + dy:q + dy + H, q + dx:q + dx + W]
    return out


def _patches_adjoint(d, H, W, dil=1):
    """Adjoint of _patches: [B,H,W,9,C] -> [B,H,W,C]."""
    B, C = d.shape[0], d.shape[-1]
    q = dil
    xp = np.zeros((B, H + 2 * q, W + 2 * q, C))
    for t in range(9):
        dy, dx = (t // 3 - 1) * dil, (t % 3 - 1) * dil
        xp[:, q + dy:q + dy + H, q + dx:q + dx + W] += d[:, :, :, t]
    return xp[:, q:q + H, q:q + W]


def _up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def _up_adjoint(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def _pool(x):
    B, H, W, C = x.shape
    v = x.reshape(B, H // 2, 2, W // 2, 2, C)
    return ((v[:, :, 0, :, 0] + v[:, :, 0, :, 1]) + (v[:, :, 1, :, 0] + v[:, :, 1, :, 1])) * 0.25


def _pool_adjoint(d):
    return _up(d) * 0.25


def forward(params, x, rnd=False):
    """x [B,R,R,4] -> (out [B,R,R], cache)."""
    p = np.asarray(params, dtype=np.float64)
    r = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    acts = {"X": r(x)}
    cache = {"inputs": {}}
    for li, (_, a, b, ca, cb, cout, o) in enumerate(LAYERS):
        xin = acts[a] if b is None else np.concatenate([acts[a], _up(acts[b])], axis=-1)
        w, bias = _conv_weights(p, li, rnd)
        P = _patches(xin, DIL[li])
        z = np.einsum("bhwtc,otc->bhwo", P, w) + bias
        acts[o] = r(np.maximum(z, 0.0))
        cache["inputs"][li] = (P, xin.shape)
        if o in POOLS:
            acts[POOLS[o]] = r(_pool(acts[o]))
    _, _, total = layout()
    hw = p[total - HEAD_C - 1:total - 1]
    hb = p[total - 1]
    out = np.einsum("bhwc,c->bhw", acts["D0"], hw) + hb
    cache["acts"] = acts
    return out, cache


def _block_mean_up(e, s):
    """Mean of e over 2^s x 2^s blocks, repeated back to full resolution."""
    B, H, W = e.shape
    b = 1 << s
    m = e.reshape(B, H // b, b, W // b, b).mean(axis=(2, 4))
    return np.repeat(np.repeat(m, b, axis=1), b, axis=2)


def loss_and_grads(params, x, target, weight, rnd=False, pyramid=0.0):
    """Loss (1/n) sum_p [w d^2 + alpha sum_s m_s(p)^2] (d = out - target, m_s
    the block means of w d over 2x2 .. 16x16 blocks) and its gradient in the
    storage layout."""
    p = np.asarray(params, dtype=np.float64)
    r = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    out, cache = forward(p, x, rnd)
    acts = cache["acts"]
    n = out.size
    diff = out - target
    ms = [_block_mean_up(weight * diff, s) for s in range(1, 5)] if pyramid else []
    loss = float((np.sum(weight * diff * diff) + pyramid * sum(np.sum(m * m) for m in ms)) / n)
    dout = 2.0 * weight * (diff + pyramid * sum(ms)) / n
    lay, _, total = layout()
    g = np.zeros(total)
    hw = p[total - HEAD_C - 1:total - 1]
    g[total - HEAD_C - 1:total - 1] = np.einsum("bhw,bhwc->c", dout, acts["D0"])
    g[total - 1] = dout.sum()
    # gradient at the pre-activation of each layer output buffer
    dz = {"D0": r(np.where(acts["D0"] > 0, dout[..., None] * hw, 0.0))}
    pending = {}  # buffer -> accumulated gradient (post-activation) from several consumers
    for li in reversed(range(len(LAYERS))):
        _, a, b, ca, cb, cout, o = LAYERS[li]
        dy = dz[o]
        P, shp = cache["inputs"][li]
        ow, ob, _, cp, kp = lay[li]
        dw = np.einsum("bhwo,bhwtc->otc", dy, P)
        gw = np.zeros((cout, kp))
        gw[:, :9 * cp] = dw.reshape(cout, 9 * cp)
        g[ow:ow + cout * kp] = gw.ravel()
        g[ob:ob + cout] = dy.sum(axis=(0, 1, 2))
        if li == 0:
            break
        w, _ = _conv_weights(p, li, rnd)
        dxin = _patches_adjoint(np.einsum("bhwo,otc->bhwtc", dy, w), shp[1], shp[2], DIL[li])
        if b is None:
            if a in ("P0", "P1", "P2"):
                # input is a pooled skip tensor: route to the skip's gradient
                src = {"P0": "E0", "P1": "E1", "P2": "E2"}[a]
                pending.setdefault(src, []).append(("pool", r(dxin)))
                e = acts[src]
                parts = pending.pop(src)
                cat = [t for k, t in parts if k == "cat"][0]
                pool = [t for k, t in parts if k == "pool"][0]
                dz[src] = r(np.where(e > 0, cat + _pool_adjoint(pool), 0.0))
            else:
                dz[a] = r(np.where(acts[a] > 0, dxin, 0.0))
        else:
            dcat = r(dxin)
            pending.setdefault(a, []).append(("cat", dcat[..., :ca]))
            dz[b] = r(np.where(acts[b] > 0, _up_adjoint(dcat[..., ca:]), 0.0))
    return loss, g, out


def init_params(seed: int = 0) -> np.ndarray:
    """A random parameter vector in the storage layout (zero padding)."""
    rng = np.random.default_rng(seed)
    lay, _, total = layout()
    p = np.zeros(total, dtype=np.float32)
    for li, (ow, ob, cout, cp, kp) in enumerate(lay):
        w = np.zeros((cout, kp), dtype=np.float32)
        blk = rng.standard_normal((cout, 9, cp)) * np.sqrt(2.0 / (9 * cp))
        w[:, :9 * cp] = blk.reshape(cout, 9 * cp)
        p[ow:ow + cout * kp] = w.ravel()
        p[ob:ob + cout] = rng.standard_normal(cout) * 0.05
    p[total - HEAD_C - 1:total] = rng.standard_normal(HEAD_C + 1) * 0.25
    return p
