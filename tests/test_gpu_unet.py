"""The warm-start network's HIP kernels against the CPU reference
(oracle/unet_ref.py, float64 with the GPU's bf16 storage points emulated):
forward output, loss and every layer's weight gradient; determinism of a
training step; the AdamW update; a short training run that must fit a fixed
batch; the coarsening kernel against numpy; and the solver's fallback to the
plain start when the network's start is worse (injected bad weights)."""
import numpy as np
import pytest

from paintop_amd import hip
from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def _dev(a):
    from paintop_amd.device import DeviceArray

    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float32))


def _case(B, R, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, R, R, 4)).astype(np.float32)
    x[..., 3] = 0
    y = rng.standard_normal((B, R, R)).astype(np.float32)
    w = rng.random((B, R, R)).astype(np.float32)
    return x, y, w


# (4, 128): 65 536 pixels at the top level, above the split-tap kernel's limit,
# so the top level's convolutions and their input gradients run on k_conv3 (its
# 32-wide K chunks, and the 16-wide remainder of the 48-channel decoder input);
# the smaller cases run every layer with 16+ input channels on k_conv3_sk
@pytest.mark.parametrize("B,R,pyr", [(2, 32, 0.0), (1, 64, 0.0), (2, 32, 0.7), (4, 128, 0.0)])
def test_forward_and_gradients_match_cpu_reference(gpu, B, R, pyr):
    from paintop_amd.device import DeviceArray

    p = U.init_params(7)
    x, y, w = _case(B, R, 5)
    net = hip.UNet(B, R, p)
    dx, dy, dw = _dev(x), _dev(y), _dev(w)
    dout = DeviceArray((B, R, R), np.float32)
    net.forward(dx.ptr, dout.ptr, B)
    out = dout.numpy()
    ref, _ = U.forward(p, x, rnd=True)
    ferr = float(np.abs(out - ref).max() / np.abs(ref).max())
    loss = net.train_step(dx.ptr, dy.ptr, dw.ptr, B, 0.0, update=False, pyramid=pyr)
    _, g, _, _, step = net.state()
    loss_ref, g_ref, _ = U.loss_and_grads(p, x, y, w, rnd=True, pyramid=pyr)
    print(f"[unet] B={B} R={R} pyramid {pyr}: forward max rel err {ferr:.2e}, loss {loss:.6f} "
          f"ref {loss_ref:.6f}", flush=True)
    assert step == 0
    assert ferr < 1e-2
    assert abs(loss - loss_ref) < 1e-2 * loss_ref
    lay, hoff, total = U.layout()
    worst = 0.0
    for li, (ow, ob, cout, cp, kp) in enumerate(lay):
        for a, b_ in ((ow, ob), (ob, ob + cout)):
            e = np.linalg.norm(g[a:b_] - g_ref[a:b_]) / max(np.linalg.norm(g_ref[a:b_]), 1e-30)
            worst = max(worst, e)
            assert e < 3e-2, (li, a == ow, e)
        # padding slots stay exactly zero
        gl = g[ow:ow + cout * kp].reshape(cout, kp)
        assert np.all(gl[:, 9 * cp:] == 0)
    eh = np.linalg.norm(g[hoff:] - g_ref[hoff:]) / np.linalg.norm(g_ref[hoff:])
    print(f"[unet] worst per-layer gradient rel err {worst:.2e}, head {eh:.2e}", flush=True)
    assert eh < 1e-2
    net.close()


def test_training_step_is_deterministic_and_adam_is_exact(gpu):
    B, R = 2, 32
    p = U.init_params(9)
    x, y, w = _case(B, R, 6)
    dx, dy, dw = _dev(x), _dev(y), _dev(w)
    a, b = hip.UNet(B, R, p), hip.UNet(B, R, p)
    la = a.train_step(dx.ptr, dy.ptr, dw.ptr, B, 1e-3, weight_decay=1e-4)
    lb = b.train_step(dx.ptr, dy.ptr, dw.ptr, B, 1e-3, weight_decay=1e-4)
    pa, ga, ma, va, sa = a.state()
    pb, gb, mb, vb, sb = b.state()
    assert la == lb and sa == sb == 1
    assert pa.tobytes() == pb.tobytes() and ga.tobytes() == gb.tobytes()
    # first AdamW step from zero moments: m = (1-b1) g, v = (1-b2) g^2
    g = ga.astype(np.float64)
    p64 = p.astype(np.float64)
    mh = g  # bias-corrected
    vh = g * g
    expect = p64 - 1e-3 * (mh / (np.sqrt(vh) + 1e-8) + 1e-4 * p64)
    assert np.allclose(pa, expect, rtol=0, atol=2e-6)
    assert np.allclose(ma, 0.1 * g, rtol=1e-6, atol=1e-12)
    # padding stays zero through the update
    lay, _, _ = U.layout()
    ow, ob, cout, cp, kp = lay[0]
    w0 = pa[ow:ow + cout * kp].reshape(cout, kp)
    assert np.all(w0[:, 9 * cp:] == 0)
    a.close()
    b.close()


def test_training_fits_a_fixed_batch(gpu):
    """A smooth target (a fixed linear map of smoothed inputs) is fitted: the
    loss falls at least 20-fold in 300 AdamW steps."""
    B, R = 4, 32
    rng = np.random.default_rng(8)
    x = rng.standard_normal((B, R, R, 4)).astype(np.float32)
    x[..., 3] = 0
    from scipy.ndimage import uniform_filter

    y = np.stack([uniform_filter(x[b, ..., 1], 5) + 0.5 * x[b, ..., 2] for b in range(B)])
    y = y.astype(np.float32)
    w = np.ones((B, R, R), np.float32)
    dx, dy, dw = _dev(x), _dev(y), _dev(w)
    net = hip.UNet(B, R, seed=3)
    l0 = net.train_step(dx.ptr, dy.ptr, dw.ptr, B, 0.0, update=False)
    for _ in range(300):
        loss = net.train_step(dx.ptr, dy.ptr, dw.ptr, B, 2e-3)
    l1 = net.train_step(dx.ptr, dy.ptr, dw.ptr, B, 0.0, update=False)
    print(f"[unet] fit: loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < l0 / 20
    net.close()


@pytest.mark.parametrize("H,W,R,f,x0,y0", [
    (70, 90, 16, 5, -7, -3),       # a thread per column, several whole cells per block
    (40, 50, 64, 1, -5, -2),       # one pixel per cell, a block spans the coarse row
    (300, 310, 32, 12, -40, 9),    # 21 cells per block: a row in two blocks, the last partial
    (130, 150, 6, 20, 3, -4),      # a wave per cell: 4 cells per block, the last block partial
    (400, 500, 2, 300, -20, -10),  # a wave per cell wider than 256 columns
])
@pytest.mark.parametrize("mode", [0, 1])
def test_coarsen_matches_numpy(gpu, mode, H, W, R, f, x0, y0):
    from paintop_amd.device import DeviceArray
    from tests.helpers import classify

    rng = np.random.default_rng(2)
    C = 3
    s = rng.random((H, W, C)).astype(np.float32)
    t = rng.random((H, W, C)).astype(np.float32)
    u = rng.random((H, W, C)).astype(np.float32)
    m = (rng.random((H, W)) > 0.4).astype(np.float32)
    ch, lam = 1, 0.25
    ds, dt, dm, du = _dev(s), _dev(t), _dev(m), _dev(u)
    din = DeviceArray((R, R, 4), np.float32)
    dtg = DeviceArray((R, R), np.float32)
    dwt = DeviceArray((R, R), np.float32)
    gpu.warm_coarsen_device(ds.ptr, dt.ptr, dm.ptr, du.ptr, W, H, C, ch, lam, x0, y0, f, R,
                            din.ptr, dtg.ptr, dwt.ptr, mode=mode)
    gi, gt, gw = din.numpy(), dtg.numpy(), dwt.numpy()
    inside = classify(m).astype(bool)
    pad = np.pad(inside, 1)
    near = pad[:-2, 1:-1] | pad[2:, 1:-1] | pad[1:-1, :-2] | pad[1:-1, 2:]
    ring = near & ~inside
    G, A, S = (a[..., ch].astype(np.float64) for a in (s, t, u))
    if mode == 0:
        P, rhs_cell = G, lam * (G - A)
    else:
        # correction problem: P = 0, rhs = the 5-point Laplacian of the source
        # (interior cells are never on the image border)
        P = np.zeros_like(G)
        rhs_cell = np.zeros_like(G)
        rhs_cell[1:-1, 1:-1] = (G[1:-1, :-2] + G[1:-1, 2:] + G[:-2, 1:-1] + G[2:, 1:-1]
                                - 4.0 * G[1:-1, 1:-1])
    for cy in range(R):
        for cx in range(R):
            ys = slice(max(0, y0 + cy * f), max(0, min(H, y0 + cy * f + f)))
            xs = slice(max(0, x0 + cx * f), max(0, min(W, x0 + cx * f + f)))
            ins, rg = inside[ys, xs], ring[ys, xs]
            ni, nr = ins.sum(), rg.sum()
            rhs = np.sum(np.where(ins, rhs_cell[ys, xs], 0.0))
            bnd = np.sum(np.where(rg, A[ys, xs] - P[ys, xs], 0.0)) / nr if nr else 0.0
            tg = np.sum(np.where(ins, S[ys, xs] - P[ys, xs], 0.0) +
                        np.where(rg, A[ys, xs] - P[ys, xs], 0.0))
            exp = [rhs, bnd, ni / f ** 2, nr / f]
            assert np.allclose(gi[cy, cx], exp, rtol=1e-5, atol=1e-6), (cy, cx, gi[cy, cx], exp)
            assert np.isclose(gt[cy, cx], tg / (ni + nr) if ni + nr else 0.0, rtol=1e-5, atol=1e-6)
            assert gw[cy, cx] == (1.0 if ni else (0.5 if nr else 0.0))


@pytest.mark.parametrize("pw,ph,pitch,f,res,x0,y0", [
    (90, 70, 128, 5, 16, -7, -3),      # grid larger than the plane, negative offsets
    (257, 200, 320, 30, 8, 3, 11),     # 8 cells per block
    (300, 60, 320, 300, 1, -20, -5),   # f > 256: one coarse cell per block
    (64, 64, 64, 1, 64, 0, 0),         # f = 1, pitch == pw
])
def test_coarsen_plane_matches_numpy(gpu, pw, ph, pitch, f, res, x0, y0):
    """The correction input the solver builds from a plane (k_coarsen_plane
    through ph_coarsen_plane_device) against numpy: per coarse cell the sum of g
    over interior cells, the interior fraction and the ring density, with the
    plane's padding rows and columns never counted."""
    from paintop_amd.device import DeviceArray

    rng = np.random.default_rng(pw * 31 + f)
    ox, oy = 5, -2
    cls = (rng.random((ph, pitch)) < 0.5).astype(np.uint8)
    cls[0, :] = cls[-1, :] = 0
    cls[:, 0] = 0
    cls[:, pw - 1:] = 0
    g = np.where(cls > 0, rng.standard_normal((ph, pitch)), np.nan).astype(np.float32)
    dg, dc = DeviceArray.from_numpy(g), DeviceArray.from_numpy(cls)
    din = DeviceArray((res, res, 4), np.float32)
    gpu.coarsen_plane_device(dg.ptr, dc.ptr, pitch, pw, ph, ox, oy, x0, y0, f, res, din.ptr)
    got = din.numpy()
    pad = np.pad(cls, 1)
    near = (pad[:-2, 1:-1] | pad[2:, 1:-1] | pad[1:-1, :-2] | pad[1:-1, 2:]) > 0
    valid = np.zeros_like(cls, dtype=bool)
    valid[1:ph - 1, 1:pw - 1] = True
    inside = (cls > 0) & valid
    ring = (cls == 0) & near & valid
    g64 = np.where(inside, g.astype(np.float64), 0.0)
    for cy in range(res):
        for cx in range(res):
            ys = slice(max(0, y0 + cy * f - oy), max(0, min(ph, y0 + cy * f + f - oy)))
            xs = slice(max(0, x0 + cx * f - ox), max(0, min(pitch, x0 + cx * f + f - ox)))
            exp = [g64[ys, xs].sum(), 0.0, inside[ys, xs].sum() / f ** 2, ring[ys, xs].sum() / f]
            assert np.allclose(got[cy, cx], exp, rtol=1e-5, atol=1e-5), (cy, cx, got[cy, cx], exp)


@pytest.mark.parametrize("pw,ph,pitch,f,res,x0,y0,lam", [
    (90, 70, 128, 5, 16, -7, -3, 0.0),      # grid larger than the plane, negative offsets
    (257, 200, 320, 30, 8, 3, 11, 0.0),     # 8 cells per block
    (300, 60, 320, 300, 1, -20, -5, 0.25),  # f > 256: one coarse cell per block
    (64, 64, 64, 1, 64, 0, 0, 0.0),         # f = 1, pitch == pw
])
def test_correction_input_matches_numpy(gpu, pw, ph, pitch, f, res, x0, y0, lam):
    """The correction rounds' input from a plane's state x + e
    (ph_correction_input_device) against numpy in float64: per coarse cell the
    sum of -r = g - (nsum(e) - (4 + lambda) e) over interior cells and of r^2,
    with e = 0 off the interior (the device forms r in FP32, as the sweeps do:
    within FP32 rounding); and with e = 0 the input equals the g-only input
    (ph_coarsen_plane_device) bit for bit."""
    from paintop_amd.device import DeviceArray

    rng = np.random.default_rng(pw * 37 + f)
    ox, oy = 5, -2
    cls = (rng.random((ph, pitch)) < 0.6).astype(np.uint8)
    cls[0, :] = cls[-1, :] = 0
    cls[:, 0] = 0
    cls[:, pw - 1:] = 0
    g = np.where(cls > 0, rng.standard_normal((ph, pitch)), np.nan).astype(np.float32)
    e = np.where(cls > 0, rng.standard_normal((ph, pitch)), 0.0).astype(np.float32)
    dg_, dc, de = DeviceArray.from_numpy(g), DeviceArray.from_numpy(cls), DeviceArray.from_numpy(e)
    din = DeviceArray((res, res, 4), np.float32)
    drsq = DeviceArray((res, res), np.float64)
    gpu.correction_input_device(de.ptr, dg_.ptr, dc.ptr, pitch, pw, ph, ox, oy, x0, y0, f, res,
                                lam, din.ptr, drsq.ptr)
    got, rsq = din.numpy(), drsq.numpy()
    e64 = e.astype(np.float64)
    ep = np.pad(e64, 1)
    nsum = ((ep[1:-1, :-2] + ep[1:-1, 2:]) + ep[:-2, 1:-1]) + ep[2:, 1:-1]
    valid = np.zeros_like(cls, dtype=bool)
    valid[1:ph - 1, 1:pw - 1] = True
    inside = (cls > 0) & valid
    r = np.where(inside, nsum - (4.0 + lam) * e64 - g.astype(np.float64), 0.0)
    pad = np.pad(cls, 1)
    near = (pad[:-2, 1:-1] | pad[2:, 1:-1] | pad[1:-1, :-2] | pad[1:-1, 2:]) > 0
    ring = (cls == 0) & near & valid
    for cy in range(res):
        for cx in range(res):
            ys = slice(max(0, y0 + cy * f - oy), max(0, min(ph, y0 + cy * f + f - oy)))
            xs = slice(max(0, x0 + cx * f - ox), max(0, min(pitch, x0 + cx * f + f - ox)))
            exp = [-r[ys, xs].sum(), 0.0, inside[ys, xs].sum() / f ** 2, ring[ys, xs].sum() / f]
            assert np.allclose(got[cy, cx], exp, rtol=1e-5, atol=1e-4), (cy, cx, got[cy, cx], exp)
            assert np.isclose(rsq[cy, cx], (r[ys, xs] ** 2).sum(), rtol=1e-4, atol=1e-5)
    # e = 0: the state's input is the g-only input, bit for bit
    dz = DeviceArray.from_numpy(np.zeros_like(e))
    din0 = DeviceArray((res, res, 4), np.float32)
    gpu.correction_input_device(dz.ptr, dg_.ptr, dc.ptr, pitch, pw, ph, ox, oy, x0, y0, f, res,
                                lam, din.ptr, drsq.ptr)
    gpu.coarsen_plane_device(dg_.ptr, dc.ptr, pitch, pw, ph, ox, oy, x0, y0, f, res, din0.ptr)
    assert din.numpy().tobytes() == din0.numpy().tobytes()


def test_bad_weights_fall_back_to_the_plain_start(gpu, monkeypatch):
    """Injected bad weights: the network's start is worse than the anchor
    start, so the solver keeps the plain start and the result is bit-identical
    to the plain solve. (These images are below the network start's size gate;
    the gate is lowered so the network runs.)"""
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import batch_case

    monkeypatch.setenv("PAINTOP_HIP_WARM_MIN_CELLS", "0")
    cases = batch_case(3, seed0=950)
    bad = U.init_params(5) * 40.0
    gpu.warm_start_set_weights(bad)
    try:
        for (w, h, s, t, m) in cases:
            o1 = DeviceArray((h, w), np.float32)
            o2 = DeviceArray((h, w), np.float32)
            d = hip.make_desc(w, h, 1, 20000, 1e-6)
            r1 = gpu.poisson_solve_device(d, s.ptr, t.ptr, m.ptr, o1.ptr)
            d.flags = hip.PH_FLAG_WARM_START_NET
            r2 = gpu.poisson_solve_device(d, s.ptr, t.ptr, m.ptr, o2.ptr)
            print(f"[unet] fallback {w}x{h}: warm rel {r2.warm_start_rel:.3e}, "
                  f"kept {r2.warm_start_used}", flush=True)
            assert r2.warm_start_used == 0 and r2.warm_start_rel > 1.0
            assert o1.numpy().tobytes() == o2.numpy().tobytes()
            assert r1.residual_history == r2.residual_history
            # FORCE_PLAIN wins over the warm-start flag
            d.flags = hip.PH_FLAG_WARM_START_NET | hip.PH_FLAG_FORCE_PLAIN
            r3 = gpu.poisson_solve_device(d, s.ptr, t.ptr, m.ptr, o2.ptr)
            assert r3.warm_start_rel == 0.0 and r3.residual_history == r1.residual_history
    finally:
        gpu.warm_start_set_weights(None)


def test_small_images_take_the_plain_start(gpu):
    """Below the learned path's size gate (2^14 interior cells per channel
    for a lone image) a channel takes the plain start: with the flag set and
    weights loaded the result and history are the plain solve's, bit for bit,
    and the network is not tried (warm_start_rel stays 0)."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray

    w = h = 96  # an ellipse of ~3.5k interior cells
    s, t, m = _fbm_case(w, h, 1, 0.5, 960, "ellipse")
    o1, o2 = DeviceArray((h, w), np.float32), DeviceArray((h, w), np.float32)
    d = hip.make_desc(w, h, 1, 20000, 1e-6)
    r1 = gpu.poisson_solve_device(d, s.ptr, t.ptr, m.ptr, o1.ptr)
    warmstart.enable()
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        r2 = gpu.poisson_solve_device(d, s.ptr, t.ptr, m.ptr, o2.ptr)
    finally:
        gpu.warm_start_set_weights(None)
    assert r2.warm_start_used == 0 and r2.warm_start_rel == 0.0
    assert o1.numpy().tobytes() == o2.numpy().tobytes()
    assert r1.residual_history == r2.residual_history


def test_warm_start_without_weights_is_an_error(gpu):
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import batch_case

    (w, h, s, t, m), = batch_case(1, seed0=960)
    o = DeviceArray((h, w), np.float32)
    d = hip.make_desc(w, h, 1, 20000, 1e-6, flags=hip.PH_FLAG_WARM_START_NET)
    with pytest.raises(hip.HipCallError) as e:
        gpu.poisson_solve_device(d, s.ptr, t.ptr, m.ptr, o.ptr)
    assert e.value.code == "E_HIP_WARM_START_UNSET" and e.value.error_class == "policy"


@pytest.mark.parametrize("n", [1024, 2048])
def test_trained_warm_start_ablation(gpu, n):
    """What the trained network gains over the method without it, on
    held-out inputs (the 4096^2 seamless-clone generator at n^2: fbm seeds
    101+c / 201+c and star seed 301 are not among the training seeds): the
    learned start is kept on every solved channel, needs fewer sweeps than the
    plain start, reaches the same tolerance, leaves outside-mask samples
    byte-identical, and agrees with the plain result within the error a 1e-6
    residual allows. The untrained network (same architecture, seeded
    initialisation) is the ablation's other arm: its start loses the race, so
    the solve equals the plain solve."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import seamless_clone_case
    from tests.helpers import classify

    src, tgt, msk = seamless_clone_case(n, channels=4)
    o_p = DeviceArray(tgt.shape, np.float32)
    o_w = DeviceArray(tgt.shape, np.float32)
    d = hip.make_desc(n, n, 4, 100000, 1e-6)
    rp = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_p.ptr)
    warmstart.enable()
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        rw = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_w.ptr)
        gpu.warm_start_set_weights(U.init_params(11))
        o_u = DeviceArray(tgt.shape, np.float32)
        ru = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_u.ptr)
    finally:
        gpu.warm_start_set_weights(None)
    print(f"[ablation] {n}^2: plain {rp.iterations} sweeps, trained net {rw.iterations} "
          f"({rp.iterations / rw.iterations:.2f}x, start rel {rw.warm_start_rel:.2e}), "
          f"untrained net {ru.iterations} (kept {ru.warm_start_used})", flush=True)
    assert rw.converged and rw.final_residual <= 1e-6
    assert rw.warm_start_used == 3  # R, G, B (alpha has r0 == 0)
    assert rw.iterations < 0.85 * rp.iterations
    assert len(rw.residual_history) == rw.iterations
    a, b, t = o_p.numpy(), o_w.numpy(), tgt.numpy()
    inside = classify(msk.numpy()).astype(bool)
    for c in range(4):
        assert b[..., c][~inside].tobytes() == t[..., c][~inside].tobytes()
    assert float(np.abs(a - b).max()) < 2e-4
    # an untrained network's start that loses the race leaves the plain solve,
    # bit for bit; one that wins has to converge in no more sweeps than the
    # plain start
    if ru.warm_start_used == 0:
        assert o_u.numpy().tobytes() == a.tobytes() and ru.residual_history == rp.residual_history
    else:
        assert ru.converged and ru.final_residual <= 1e-6
        assert ru.iterations <= rp.iterations


@pytest.mark.parametrize("scale", [0.05, 0.2, 0.5, 1.0])
def test_degraded_network_never_costs_sweeps(gpu, scale):
    """The trained weights with multiplicative noise of growing scale (seeded):
    whatever the start's quality, a start the race keeps needs no more sweeps
    than the plain start on its channel, and a start it drops leaves the plain
    solve, bit for bit. The race is judged on the FP64 residual of both members
    at the same sweep and extended while the warm lead is under 4x."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import seamless_clone_case

    n = 1024
    src, tgt, msk = seamless_clone_case(n, channels=3)
    w0 = warmstart.load_weights()
    rng = np.random.default_rng(7000 + int(scale * 100))
    noisy = (w0 * (1.0 + scale * rng.standard_normal(w0.shape))).astype(w0.dtype)
    d = hip.make_desc(n, n, 3, 100000, 1e-6)
    o_p = DeviceArray(tgt.shape, np.float32)
    rp = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_p.ptr)
    reps_p = []
    for c in range(3):
        dc = hip.make_desc(n, n, 1, 100000, 1e-6)
        one = [DeviceArray.from_numpy(np.ascontiguousarray(v.numpy()[..., c]))
               for v in (src, tgt)]
        oc = DeviceArray((n, n), np.float32)
        reps_p.append((one, oc, gpu.poisson_solve_device(dc, one[0].ptr, one[1].ptr, msk.ptr,
                                                          oc.ptr)))
    gpu.warm_start_set_weights(noisy)
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        o_w = DeviceArray(tgt.shape, np.float32)
        rw = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_w.ptr)
        per = []
        for c, (one, oc, r1) in enumerate(reps_p):
            dc = hip.make_desc(n, n, 1, 100000, 1e-6, flags=hip.PH_FLAG_WARM_START_NET)
            ow = DeviceArray((n, n), np.float32)
            per.append((r1, gpu.poisson_solve_device(dc, one[0].ptr, one[1].ptr, msk.ptr, ow.ptr),
                        oc, ow))
    finally:
        gpu.warm_start_set_weights(None)
    print(f"[degraded] noise {scale}: plain {rp.iterations} sweeps, net {rw.iterations} "
          f"(kept {rw.warm_start_used}, start rel {rw.warm_start_rel:.2e}); per channel "
          + ", ".join(f"{r1.iterations}->{r2.iterations}({r2.warm_start_used})"
                      for r1, r2, _a, _b in per), flush=True)
    assert rw.converged and rw.final_residual <= 1e-6
    assert rw.iterations <= rp.iterations
    for r1, r2, oc, ow in per:
        assert r2.converged and r2.final_residual <= 1e-6
        if r2.warm_start_used:
            assert r2.iterations <= r1.iterations
        else:
            assert ow.numpy().tobytes() == oc.numpy().tobytes()
            assert r2.residual_history == r1.residual_history


def test_coarse_corrections_cut_sweeps(gpu, monkeypatch):
    """The learned coarse corrections (the network solving Laplacian(e) = g on
    the coarse grid after each cycle) against the same learned start without
    them (PAINTOP_HIP_CORRECTIONS=0, the warm start's own ramp): corrections
    are kept, the solve needs fewer sweeps, reaches the tolerance, leaves
    outside-mask samples byte-identical and agrees with the plain result
    within the error a 1e-6 residual allows. Held-out input (the seamless-clone
    generator at 2048^2)."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import seamless_clone_case
    from tests.helpers import classify

    n = 2048
    src, tgt, msk = seamless_clone_case(n, channels=4)
    o_p, o_c, o_n = (DeviceArray(tgt.shape, np.float32) for _ in range(3))
    d = hip.make_desc(n, n, 4, 100000, 1e-6)
    rp = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_p.ptr)
    warmstart.enable()
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        rc = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_c.ptr)
        monkeypatch.setenv("PAINTOP_HIP_CORRECTIONS", "0")
        rn = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_n.ptr)
    finally:
        gpu.warm_start_set_weights(None)
    print(f"[corrections] {n}^2: plain {rp.iterations} sweeps, learned start only "
          f"{rn.iterations}, with corrections {rc.iterations} ({rc.coarse_corrections} kept)",
          flush=True)
    assert rn.coarse_corrections == 0
    assert rc.coarse_corrections >= 3
    assert rc.converged and rc.final_residual <= 1e-6 and rn.converged
    assert rc.iterations < 0.6 * rn.iterations
    assert len(rc.residual_history) == rc.iterations
    a, b, t = o_p.numpy(), o_c.numpy(), tgt.numpy()
    inside = classify(msk.numpy()).astype(bool)
    for c in range(4):
        assert b[..., c][~inside].tobytes() == t[..., c][~inside].tobytes()
    assert float(np.abs(a - b).max()) < 2e-4


def _fbm_case(W, H, C, freq, seed, mask_kind):
    from paintop_amd.device import DeviceArray, synchronize

    src = DeviceArray((H, W, C), np.float32)
    tgt = DeviceArray((H, W, C), np.float32)
    msk = DeviceArray((H, W), np.float32)
    for c in range(C):
        hip.fbm_device(src.ptr, W, H, freq, seed + c, 6, 2.0, 0.5, channel_stride=C,
                       channel_offset=c, scale=0.5, bias=0.5)
        hip.fbm_device(tgt.ptr, W, H, freq, seed + 50 + c, 6, 2.0, 0.5, channel_stride=C,
                       channel_offset=c, scale=0.5, bias=0.5)
    if mask_kind == "ellipse":
        hip.ellipse_mask_device(msk.ptr, W, H, 0.52 * W, 0.47 * H, 0.36 * W, 0.38 * H)
    else:
        pts = [(0.12 * W, 0.20 * H), (0.55 * W, 0.08 * H), (0.90 * W, 0.35 * H),
               (0.80 * W, 0.88 * H), (0.40 * W, 0.70 * H), (0.15 * W, 0.85 * H)]
        hip.polygon_mask_device(msk.ptr, W, H, pts)
    synchronize()
    return src, tgt, msk


@pytest.mark.parametrize("W,H,C,freq,mask_kind", [
    (2560, 1600, 3, 0.2, "ellipse"),
    (3000, 3000, 1, 0.08, "polygon"),
])
def test_learned_path_on_other_shapes(gpu, W, H, C, freq, mask_kind):
    """Held-out shapes and layouts for the learned path (the network was trained
    on square crops of 256-1024 px; corrections run at these sizes): a
    non-square RGB blend with an ellipse mask and a single-channel blend with a
    concave polygon. The learned solve converges in fewer sweeps than the plain
    one, leaves outside-mask samples byte-identical and agrees with the plain
    result within the error a 1e-6 residual allows."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray
    from tests.helpers import classify

    src, tgt, msk = _fbm_case(W, H, C, freq, 700, mask_kind)
    o_p, o_w = DeviceArray(tgt.shape, np.float32), DeviceArray(tgt.shape, np.float32)
    d = hip.make_desc(W, H, C, 100000, 1e-6)
    rp = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_p.ptr)
    warmstart.enable()
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        rw = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_w.ptr)
    finally:
        gpu.warm_start_set_weights(None)
    print(f"[shapes] {W}x{H}x{C} {mask_kind}: plain {rp.iterations} sweeps, learned "
          f"{rw.iterations} ({rw.coarse_corrections} corrections, kept {rw.warm_start_used})",
          flush=True)
    assert rp.converged and rw.converged and rw.final_residual <= 1e-6
    assert rw.iterations < rp.iterations
    a, b, t = o_p.numpy(), o_w.numpy(), tgt.numpy()
    inside = classify(msk.numpy()).astype(bool)
    for c in range(C):
        assert b[..., c][~inside].tobytes() == t[..., c][~inside].tobytes()
    assert float(np.abs(a - b).max()) < 2e-4


def test_corrections_in_a_batched_call(gpu):
    """Two large blends of different sizes and masks in one batched call with
    the learned path: their warm starts, races and correction cycles run in the
    same sweep launches but end at different sweeps. Each converges, keeps its
    outside-mask samples and agrees with its own plain solve within the error a
    1e-6 residual allows."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray
    from tests.helpers import classify

    cases = [_fbm_case(2048, 2048, 3, 0.125, 810, "polygon"),
             _fbm_case(1800, 1400, 1, 0.2, 820, "ellipse")]
    dims = [(2048, 2048, 3), (1800, 1400, 1)]
    d = hip.make_desc(0, 0, 1, 100000, 1e-6)
    plain, learned = [], []
    for (w, h, c), (s_, t_, m_) in zip(dims, cases):
        o = DeviceArray(t_.shape, np.float32)
        r = gpu.poisson_solve_device(hip.make_desc(w, h, c, 100000, 1e-6), s_.ptr, t_.ptr,
                                     m_.ptr, o.ptr)
        plain.append((o, r))
    outs = [DeviceArray(t_.shape, np.float32) for (_s, t_, _m) in cases]
    items = [(w, h, c, s_.ptr, t_.ptr, m_.ptr, o.ptr)
             for (w, h, c), (s_, t_, m_), o in zip(dims, cases, outs)]
    warmstart.enable()
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        reps = gpu.poisson_solve_batch_device(d, items)
    finally:
        gpu.warm_start_set_weights(None)
    for (w, h, c), (s_, t_, m_), o, r, (op, rp) in zip(dims, cases, outs, reps, plain):
        print(f"[batch-learned] {w}x{h}x{c}: plain {rp.iterations}, learned {r.iterations} "
              f"({r.coarse_corrections} corrections, kept {r.warm_start_used})", flush=True)
        assert r.converged and r.final_residual <= 1e-6 and r.coarse_corrections > 0
        assert r.iterations < rp.iterations
        a, b, t = op.numpy(), o.numpy(), t_.numpy()
        inside = classify(m_.numpy()).astype(bool)
        for ch in range(c):
            bc = b[..., ch] if c > 1 else b
            tc = t[..., ch] if c > 1 else t
            assert bc[~inside].tobytes() == tc[~inside].tobytes()
        assert float(np.abs(a - b).max()) < 2e-4


def test_batch_gate_is_relative_to_the_largest_image(gpu):
    """In one batched call with the learned path, an image far smaller than the
    call's largest (below max(2^14, largest / 8) interior cells per channel)
    takes the plain start: its output, history and report equal the plain
    batched call's, bit for bit, while the large image keeps its warm start and
    corrections. The same small image alone passes the gate (2^14) and takes
    the learned path."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray

    # interior cells per channel: ~473k, ~24.6k, ~25.9k (gate in the batch: ~59k)
    dims = [(1024, 1024, 3), (256, 224, 1), (224, 256, 2)]
    cases = [_fbm_case(1024, 1024, 3, 0.125, 830, "polygon"),
             _fbm_case(256, 224, 1, 0.3, 840, "ellipse"),
             _fbm_case(224, 256, 2, 0.3, 850, "polygon")]
    d = hip.make_desc(0, 0, 1, 100000, 1e-6)

    def batch(flags):
        outs = [DeviceArray(t_.shape, np.float32) for (_s, t_, _m) in cases]
        items = [(w, h, c, s_.ptr, t_.ptr, m_.ptr, o.ptr)
                 for (w, h, c), (s_, t_, m_), o in zip(dims, cases, outs)]
        d.flags = flags
        return outs, gpu.poisson_solve_batch_device(d, items)

    outs_p, reps_p = batch(0)
    warmstart.enable()
    try:
        outs_l, reps_l = batch(hip.PH_FLAG_WARM_START_NET)
        w, h, c = dims[1]
        s_, t_, m_ = cases[1]
        o = DeviceArray(t_.shape, np.float32)
        alone = gpu.poisson_solve_device(
            hip.make_desc(w, h, c, 100000, 1e-6, flags=hip.PH_FLAG_WARM_START_NET),
            s_.ptr, t_.ptr, m_.ptr, o.ptr)
    finally:
        gpu.warm_start_set_weights(None)
    for k, ((w, h, c), rp, rl) in enumerate(zip(dims, reps_p, reps_l)):
        print(f"[gate] {w}x{h}x{c}: plain {rp.iterations}, learned {rl.iterations} "
              f"(kept {rl.warm_start_used}, corrections {rl.coarse_corrections})", flush=True)
    big = reps_l[0]
    assert big.converged and big.warm_start_used > 0 and big.coarse_corrections > 0
    assert big.iterations < reps_p[0].iterations
    for k in (1, 2):
        assert reps_l[k].warm_start_used == 0 and reps_l[k].warm_start_rel == 0.0, k
        assert outs_l[k].numpy().tobytes() == outs_p[k].numpy().tobytes(), k
        assert reps_l[k].residual_history == reps_p[k].residual_history, k
        assert reps_l[k].iterations == reps_p[k].iterations, k
    print(f"[gate] 256x224x1 alone: learned {alone.iterations} (kept {alone.warm_start_used})",
          flush=True)
    assert alone.converged and alone.warm_start_rel > 0.0


def test_corrections_taken_back(gpu, monkeypatch):
    """The take-back path of the correction rounds: with the largest allowed
    residual jump set below 1 every correction raises the residual too much
    (a good one leaves it within ~1.6x), so each is taken back in the same
    round (x -= the same step, FP64 residual again, in one launch for all
    channels), the second in a row ends the channel's cycles, none is counted,
    and the solve still converges to the tolerance with the outside-mask
    samples kept, within the error a 1e-6 residual allows of the plain solve."""
    from paintop_amd import warmstart
    from paintop_amd.device import DeviceArray
    from tests.helpers import classify

    w, h, c = 1024, 1024, 3
    s_, t_, m_ = _fbm_case(w, h, c, 0.125, 860, "polygon")
    d = hip.make_desc(w, h, c, 100000, 1e-6)
    op = DeviceArray(t_.shape, np.float32)
    rp = gpu.poisson_solve_device(d, s_.ptr, t_.ptr, m_.ptr, op.ptr)
    ol, oc = DeviceArray(t_.shape, np.float32), DeviceArray(t_.shape, np.float32)
    warmstart.enable()
    try:
        d.flags = hip.PH_FLAG_WARM_START_NET
        rc = gpu.poisson_solve_device(d, s_.ptr, t_.ptr, m_.ptr, oc.ptr)
        monkeypatch.setenv("PAINTOP_HIP_CORR_MAX_JUMP", "0.5")
        rl = gpu.poisson_solve_device(d, s_.ptr, t_.ptr, m_.ptr, ol.ptr)
    finally:
        gpu.warm_start_set_weights(None)
    print(f"[take-back] plain {rp.iterations}, corrected {rc.iterations} "
          f"({rc.coarse_corrections} kept), taken back {rl.iterations} "
          f"({rl.coarse_corrections} kept, start kept {rl.warm_start_used})", flush=True)
    assert rc.coarse_corrections > 0
    assert rl.warm_start_used > 0 and rl.coarse_corrections == 0
    assert rl.converged and rl.final_residual <= 1e-6
    assert rl.iterations > rc.iterations
    a, b, t = op.numpy(), ol.numpy(), t_.numpy()
    inside = classify(m_.numpy()).astype(bool)
    for ch in range(c):
        assert b[..., ch][~inside].tobytes() == t[..., ch][~inside].tobytes()
    assert float(np.abs(a - b).max()) < 2e-4


def test_blend_split_into_mask_components(gpu):
    """components.py: the blend solved per 4-connected mask component in one
    batched call equals the whole-mask solve (both at 1e-9: within 1e-6), keeps
    the outside-mask samples, and every component meets the tolerance on its
    own (so the whole mask meets it: ||r||^2 = sum ||r_k||^2). Through the op
    (PoissonBlend(split_components=True)) the report is the worst component's."""
    from paintop_amd import components as C
    from paintop_amd import ops
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import spots_case
    from paintop_amd.values import image, mask

    n = 1024
    src, tgt, m = spots_case(n, 12, 3, seed=602)
    S, T = src.numpy(), tgt.numpy()
    dm = DeviceArray.from_numpy(m)
    out = DeviceArray(tgt.shape, np.float32)
    gpu.poisson_solve_device(hip.make_desc(n, n, 3, 100000, 1e-9), src.ptr, tgt.ptr, dm.ptr,
                             out.ptr)
    cand, worst, per = C.solve_components(hip.make_desc(0, 0, 1, 100000, 1e-9), S, T, m)
    assert len(per) == 12 and all(r.converged and r.final_residual <= 1e-9 for _c, r in per)
    assert worst.iterations == max(r.iterations for _c, r in per)
    inside = C.interior_cells(m)
    assert cand[~inside].tobytes() == T[~inside].tobytes()
    assert float(np.abs(cand - out.numpy()).max()) < 1e-6
    op = ops.PoissonBlend(split_components=True)
    res = op.compute({"source": image(n, n, S.reshape(-1), "Rgb"),
                      "target": image(n, n, T.reshape(-1), "Rgb"), "mask": mask(n, n, m)},
                     {"tolerance": 1e-6, "max_iterations": 20000})
    solver = res["report"].as_report().solver
    assert solver.converged and len(op.last_components) == 12
    assert solver.iterations == max(r.iterations for _c, r in op.last_components)


def test_start_field_race_keeps_the_better_start(gpu):
    """ph_poisson_solve_start_device races the caller's start against the
    anchor start: a start at the solution wins and ends within a few dozen
    sweeps; a start far off loses, and the result is then the plain solve, bit
    for bit (the plain member of the race is unperturbed by it)."""
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import seamless_clone_case
    from tests.helpers import classify

    n = 512
    src, tgt, msk = seamless_clone_case(n, channels=1)
    o_p = DeviceArray(tgt.shape, np.float32)
    o_t = DeviceArray(tgt.shape, np.float32)
    o_s = DeviceArray(tgt.shape, np.float32)
    d = hip.make_desc(n, n, 1, 100000, 1e-6)
    rp = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_p.ptr)
    rt = gpu.poisson_solve_device(hip.make_desc(n, n, 1, 100000, 1e-10), src.ptr, tgt.ptr,
                                  msk.ptr, o_t.ptr)
    good = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_s.ptr, d_start=o_t.ptr)
    print(f"[race] plain {rp.iterations} sweeps; start at the 1e-10 solution: "
          f"{good.iterations} sweeps (kept {good.warm_start_used})", flush=True)
    assert good.warm_start_used == 1 and good.converged and good.iterations < 64
    assert float(np.abs(o_s.numpy() - o_t.numpy()).max()) < 1e-5
    inside = classify(msk.numpy()).astype(bool)
    bad_start = np.where(inside, tgt.numpy() + 1000.0, tgt.numpy()).astype(np.float32)
    dbad = DeviceArray.from_numpy(bad_start)
    bad = gpu.poisson_solve_device(d, src.ptr, tgt.ptr, msk.ptr, o_s.ptr, d_start=dbad.ptr)
    print(f"[race] start far off: {bad.iterations} sweeps (kept {bad.warm_start_used}, "
          f"start rel {bad.warm_start_rel:.2e})", flush=True)
    assert bad.warm_start_used == 0
    assert o_s.numpy().tobytes() == o_p.numpy().tobytes()
    assert bad.residual_history == rp.residual_history and bad.iterations == rp.iterations


def test_gather_dihedral_matches_numpy(gpu):
    from paintop_amd.device import DeviceArray

    rng = np.random.default_rng(21)
    n, R, B = 11, 16, 8
    X = rng.random((n, R, R, 4)).astype(np.float32)
    Y = rng.random((n, R, R)).astype(np.float32)
    W = rng.random((n, R, R)).astype(np.float32)
    dX, dY, dW = _dev(X), _dev(Y), _dev(W)
    bX, bY, bW = (DeviceArray(s, np.float32) for s in ((B, R, R, 4), (B, R, R), (B, R, R)))
    idx = rng.integers(0, n, B)
    modes = np.arange(B)
    gpu.unet_gather_device(dX.ptr, dY.ptr, dW.ptr, n, idx, modes, R, bX.ptr, bY.ptr, bW.ptr)

    def mapped(a, m):  # out[y, x] = a[sy, sx]
        yy, xx = np.mgrid[0:R, 0:R]
        sy, sx = (xx, yy) if m & 4 else (yy, xx)
        sx = R - 1 - sx if m & 1 else sx
        sy = R - 1 - sy if m & 2 else sy
        return a[sy, sx]

    gx, gy, gw = bX.numpy(), bY.numpy(), bW.numpy()
    for j in range(B):
        assert np.array_equal(gy[j], mapped(Y[idx[j]], modes[j]))
        assert np.array_equal(gw[j], mapped(W[idx[j]], modes[j]))
        assert np.array_equal(gx[j], mapped(X[idx[j]], modes[j]))
    # the 8 maps are distinct
    assert len({mapped(np.arange(R * R).reshape(R, R), m).tobytes() for m in range(8)}) == 8
    with pytest.raises(hip.HipCallError):
        gpu.unet_gather_device(dX.ptr, dY.ptr, dW.ptr, n, [n], None, R, bX.ptr, bY.ptr, bW.ptr)
