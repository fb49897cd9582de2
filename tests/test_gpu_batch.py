"""The batched entry point: every image of a batch gets exactly the result
(bits and report) it gets when solved alone, and the 512-instance uneven
batch converges per instance with outside-mask samples untouched."""
import numpy as np
import pytest

from paintop_amd import hip
from tests.helpers import classify

pytestmark = pytest.mark.gpu


def _items(cases):
    from paintop_amd.device import DeviceArray

    items, outs = [], []
    for (w, h, s, t, m) in cases:
        o = DeviceArray((h, w), np.float32)
        outs.append(o)
        items.append((w, h, 1, s.ptr, t.ptr, m.ptr, o.ptr))
    return items, outs


def test_batch_equals_solo(gpu):
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import batch_case

    cases = batch_case(14, seed0=900)
    # add an image with no interior and one with a flat source/target (r0 == 0)
    w, h = 40, 30
    flat = DeviceArray.from_numpy(np.full((h, w), 0.5, np.float32))
    ones = DeviceArray.from_numpy(np.ones((h, w), np.float32))
    zeros = DeviceArray.from_numpy(np.zeros((h, w), np.float32))
    cases.append((w, h, flat, flat, ones))
    cases.append((w, h, flat, ones, zeros))
    desc = hip.make_desc(0, 0, 1, 20000, 1e-6)
    items, outs = _items(cases)
    res = gpu.poisson_solve_batch_device(desc, items)
    for k, (w, h, s, t, m) in enumerate(cases):
        o = DeviceArray((h, w), np.float32)
        r = gpu.poisson_solve_device(gpu.make_desc(w, h, 1, 20000, 1e-6), s.ptr, t.ptr, m.ptr,
                                     o.ptr)
        assert outs[k].numpy().tobytes() == o.numpy().tobytes(), k
        assert res[k].iterations == r.iterations and res[k].stop_reason == r.stop_reason, k
        assert res[k].residual_history == r.residual_history, k
        assert res[k].final_residual == r.final_residual, k
    assert res[-2].iterations == 0 and res[-1].iterations == 0


def _multichannel_case(W, H, C, seed):
    from paintop_amd.device import DeviceArray, synchronize

    src = DeviceArray((H, W, C), np.float32)
    tgt = DeviceArray((H, W, C), np.float32)
    msk = DeviceArray((H, W), np.float32)
    for c in range(C):
        hip.fbm_device(src.ptr, W, H, 0.2, seed + c, 6, 2.0, 0.5, channel_stride=C,
                       channel_offset=c, scale=0.5, bias=0.5)
        hip.fbm_device(tgt.ptr, W, H, 0.2, seed + 50 + c, 6, 2.0, 0.5, channel_stride=C,
                       channel_offset=c, scale=0.5, bias=0.5)
    hip.ellipse_mask_device(msk.ptr, W, H, 0.5 * W, 0.45 * H, 0.35 * W, 0.3 * H)
    synchronize()
    return src, tgt, msk


def test_mixed_channel_counts_equal_solo(gpu):
    """Images of 1, 3, 5 and 4 channels in one call: the per-image assembly and
    write-back then take their general form (any channel count), solo solves
    their per-count forms (1..4); both give the same bits and reports, and a
    channel with a flat source and target (r0 = 0) keeps the target's bits."""
    from paintop_amd.device import DeviceArray

    dims = [(96, 80, 1), (120, 100, 3), (64, 72, 5), (80, 64, 4)]
    cases = [_multichannel_case(w, h, c, 700 + 10 * i) for i, (w, h, c) in enumerate(dims)]
    # a flat last channel in the 4-channel image
    s4, t4, _m = cases[3]
    a, b = s4.numpy(), t4.numpy()
    a[..., 3] = 0.25
    b[..., 3] = 0.25
    cases[3] = (DeviceArray.from_numpy(a), DeviceArray.from_numpy(b), cases[3][2])
    outs = [DeviceArray(t_.shape, np.float32) for (_s, t_, _m) in cases]
    items = [(w, h, c, s_.ptr, t_.ptr, m_.ptr, o.ptr)
             for (w, h, c), (s_, t_, m_), o in zip(dims, cases, outs)]
    res = gpu.poisson_solve_batch_device(hip.make_desc(0, 0, 1, 20000, 1e-6), items)
    for k, ((w, h, c), (s_, t_, m_)) in enumerate(zip(dims, cases)):
        o = DeviceArray(t_.shape, np.float32)
        r = gpu.poisson_solve_device(hip.make_desc(w, h, c, 20000, 1e-6), s_.ptr, t_.ptr,
                                     m_.ptr, o.ptr)
        assert outs[k].numpy().tobytes() == o.numpy().tobytes(), k
        assert res[k].iterations == r.iterations and res[k].worst_channel == r.worst_channel, k
        assert res[k].residual_history == r.residual_history, k
        assert res[k].final_residual == r.final_residual, k
        assert r.converged, k
    flat = outs[3].numpy()[..., 3]
    assert flat.tobytes() == cases[3][1].numpy()[..., 3].tobytes()


def test_uneven_batch_of_512(gpu):
    from paintop_amd.inputs import batch_case

    cases = batch_case(512, seed0=500)
    desc = hip.make_desc(0, 0, 1, 100000, 1e-6)
    items, outs = _items(cases)
    res = gpu.poisson_solve_batch_device(desc, items, want_history=False)
    its = np.array([r.iterations for r in res])
    px = sum(w * h for (w, h, *_r) in cases)
    print(f"[batch] 512 instances, {px / 1e6:.1f} MP: sweeps min {its.min()} median "
          f"{int(np.median(its))} max {its.max()}, launches {res[0].launches}", flush=True)
    assert all(r.converged and r.final_residual <= 1e-6 for r in res)
    for k in (0, 1, 100, 511):
        w, h, s, t, m = cases[k]
        inside = classify(m.numpy()).astype(bool)
        assert outs[k].numpy()[~inside].tobytes() == t.numpy()[~inside].tobytes()
