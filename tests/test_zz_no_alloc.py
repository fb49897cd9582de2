"""The allocation-free entry point (ph_poisson_solve_device_noalloc) against
ph_poisson_solve_device on the same device inputs: bit-identical candidates and
identical reports, on two calls in a row (the second smaller, so it runs on the
first call's leftover workspace contents), then an ordinary call again, a
rejected oversized solve and rejected flags. Every call goes through the C ABI.
"""
import numpy as np
import pytest

from paintop_amd import hip

pytestmark = pytest.mark.gpu


def _case(w, h, c, seed):
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import star_polygon

    src = DeviceArray((h, w, c), np.float32)
    tgt = DeviceArray((h, w, c), np.float32)
    msk = DeviceArray((h, w), np.float32)
    for k in range(c):
        hip.fbm_device(src.ptr, w, h, 0.5, seed + k, 6, channel_stride=c, channel_offset=k,
                       scale=0.5, bias=0.5)
        hip.fbm_device(tgt.ptr, w, h, 0.5, seed + 50 + k, 6, channel_stride=c,
                       channel_offset=k, scale=0.5, bias=0.5)
    verts = star_polygon(w / 2, h / 2, 0.45 * min(w, h), 0.5, 6, 0.1, seed)
    hip.polygon_mask_device(msk.ptr, w, h, verts)
    return src, tgt, msk


def _both(w, h, c, seed, tol=1e-6):
    from paintop_amd.device import DeviceArray

    src, tgt, msk = _case(w, h, c, seed)
    desc = hip.make_desc(w, h, c, 20000, tol)
    o_ref = DeviceArray((h, w, c), np.float32)
    o_na = DeviceArray((h, w, c), np.float32)
    print(f"[noalloc] {w}x{h}x{c} seed {seed}", flush=True)
    r_ref = hip.poisson_solve_device(desc, src.ptr, tgt.ptr, msk.ptr, o_ref.ptr)
    r_na = hip.poisson_solve_device_noalloc(desc, src.ptr, tgt.ptr, msk.ptr, o_na.ptr)
    a, b = o_ref.numpy(), o_na.numpy()
    for x in (src, tgt, msk, o_ref, o_na):
        x.free()
    return a, b, r_ref, r_na


def test_noalloc_matches_the_allocating_entry_point(gpu):
    for (w, h, c, seed) in ((480, 400, 3, 11), (200, 150, 1, 23)):
        a, b, ra, rb = _both(w, h, c, seed)
        assert ra.converged and ra.iterations > 0
        assert a.tobytes() == b.tobytes()
        assert (ra.iterations, ra.stop_reason, ra.final_residual, ra.worst_channel) == \
            (rb.iterations, rb.stop_reason, rb.final_residual, rb.worst_channel)
        assert ra.residual_history == rb.residual_history
        print(f"[noalloc] {w}x{h}x{c}: {ra.iterations} sweeps, bit-identical", flush=True)
    # the allocating path still works after the allocation-free calls
    a, b, ra, rb = _both(96, 96, 2, 31)
    assert a.tobytes() == b.tobytes() and ra.residual_history == rb.residual_history


def test_noalloc_rejects_oversized_solves_and_flags(gpu):
    from paintop_amd.device import DeviceArray

    n = 2048  # a 2048^2 RGBA solve needs far more than the fixed workspace
    src, tgt, msk = _case(n, n, 4, 41)
    out = DeviceArray((n, n, 4), np.float32)
    with pytest.raises(hip.HipCallError) as e:
        hip.poisson_solve_device_noalloc(hip.make_desc(n, n, 4, 1000, 1e-6), src.ptr, tgt.ptr,
                                         msk.ptr, out.ptr)
    assert e.value.code == "E_HIP_SCRATCH_LIMIT" and e.value.error_class == "policy"
    for flag in (hip.PH_FLAG_WARM_START_NET, hip.PH_FLAG_TIMING):
        with pytest.raises(hip.HipCallError) as e:
            hip.poisson_solve_device_noalloc(hip.make_desc(64, 64, 1, 100, 1e-6, flags=flag),
                                             src.ptr, tgt.ptr, msk.ptr, out.ptr)
        assert e.value.code == "E_HIP_NOALLOC_FLAG" and e.value.error_class == "schema"
    for x in (src, tgt, msk, out):
        x.free()
