"""GPU tests of assert.no_change_outside_mask@1's device scan
(ph_assert_no_change_outside_mask_device) against the CPU oracle
(oracle/outside_scan.c, a restatement of crates/paintop-cpu/src/assert.rs:405-446):
bit-exact in the encoded space (integer counts, exact maxima and row-major
locations), within one f32 ulp per decoded sample in the decoded-linear space;
the reference's fixtures (assert/tests.rs:148-306) through the host op; and
the solver's candidate against its target at bench size."""
import numpy as np
import pytest

from oracle import oracle as O
from paintop_amd import hip, ops
from paintop_amd.errors import PaintopError
from paintop_amd.values import Extent, ImageDescriptor, MaskDescriptor, ResourceValue

pytestmark = pytest.mark.gpu


def _dev(a):
    from paintop_amd.device import DeviceArray
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float32))


def _scan(b, a, m, **kw):
    h, w = m.shape
    c = 1 if b.ndim == 2 else b.shape[2]
    db, da, dm = _dev(b), _dev(a), _dev(m)
    try:
        return hip.assert_no_change_outside_mask_device(db.ptr, da.ptr, dm.ptr, w, h, c, **kw)
    finally:
        for x in (db, da, dm):
            x.free()


def _case(rng, h, w, c, p_out=0.6, p_change=0.3):
    b = rng.random((h, w, c)).astype(np.float32)
    a = b.copy()
    ch = rng.random((h, w, c)) < p_change
    a[ch] += rng.choice([-0.5, -0.2, -0.05, 0.01, 0.05, 0.3], size=ch.sum()).astype(np.float32)
    m = np.where(rng.random((h, w)) < p_out, 0.0, rng.random((h, w))).astype(np.float32)
    return b, a, m


@pytest.mark.parametrize("h,w,c,thr,eps", [
    (1, 1, 1, 0.0, 0.0), (3, 5, 2, 0.1, 0.0), (131, 257, 3, 0.04, 0.25),
    (700, 1000, 4, 0.1, 0.5), (2048, 2049, 1, 0.0, 0.0),
])
def test_device_scan_matches_oracle_encoded(gpu, h, w, c, thr, eps):
    rng = np.random.default_rng(h * 7919 + w)
    b, a, m = _case(rng, h, w, c)
    if h * w > 10:
        a[h // 2, w // 3, 0] = np.nan       # NaN never counts
    bb, aa = (x[..., 0] if c == 1 else x for x in (b, a))
    got = _scan(bb, aa, m, outside_threshold=thr, coverage_epsilon=eps)
    exp = O.outside_scan(bb, aa, m, outside_threshold=thr, coverage_epsilon=eps)
    assert got == exp


def test_device_scan_matches_oracle_decoded(gpu):
    """decoded-linear: every change is 0 or >= 0.01 in encoded terms, so the
    leak set is the same under a one-ulp decode difference; the maxima agree to
    one ulp of the decoded values."""
    rng = np.random.default_rng(5)
    b, a, m = _case(rng, 600, 900, 3)
    a = np.clip(a, 0.0, 1.0)
    got = _scan(b, a, m, decode_srgb=True)
    exp = O.outside_scan(b, a, m, decode=True)
    assert got["changed_outside"] == exp["changed_outside"]
    assert got["locations"] == exp["locations"] and got["worst_pixel"] == exp["worst_pixel"]
    assert abs(got["max_outside"] - exp["max_outside"]) <= 2.5e-7 * max(exp["max_outside"], 1e-30)


def test_location_cap_across_blocks(gpu):
    """More than 256 leaks spread over several scan blocks (4096 pixels each):
    the list holds exactly the first 256 in row-major order and the count is
    exact."""
    h, w = 300, 200
    b = np.zeros((h, w), np.float32)
    a = b.copy()
    m = np.ones((h, w), np.float32)
    idx = np.arange(0, h * w, 97)              # one leak every 97 pixels
    m.flat[idx] = 0.0
    a.flat[idx] = np.linspace(0.1, 0.9, idx.size, dtype=np.float32)
    got = _scan(b, a, m)
    assert got == O.outside_scan(b, a, m)
    assert got["changed_outside"] == idx.size and len(got["locations"]) == 256
    assert got["worst_pixel"] == (int(idx[-1]) % w, int(idx[-1]) // w)


def test_reference_fixtures_through_the_op(gpu):
    """assert/tests.rs:148-306 through ops.NoChangeOutsideMask (the scan on the
    GPU)."""
    op = ops.NoChangeOutsideMask()
    img = lambda w, h, s, col="LinearSrgb": ResourceValue(  # noqa: E731
        ImageDescriptor(Extent(w, h), "Gray", col), 1, s)
    msk = lambda w, h, s: ResourceValue(MaskDescriptor(Extent(w, h)), 1, s)  # noqa: E731
    out = lambda ins, p: op.compute(ins, p)["report"].report.assertion  # noqa: E731

    before = np.full(16, 0.2, np.float32)
    after = before.copy()
    mk = np.zeros(16, np.float32)
    for i in (0, 1, 4, 5):
        after[i], mk[i] = 0.9, 1.0
    o = out({"before": img(4, 4, before), "after": img(4, 4, after), "allowed": msk(4, 4, mk)},
            None)
    assert o.passed and o.changed_pixels_outside == 0 and o.max_abs_delta_outside == 0.0
    assert o.worst_pixel is None and o.locations == []
    after[15] = 0.5
    o = out({"before": img(4, 4, before), "after": img(4, 4, after), "allowed": msk(4, 4, mk)},
            {"comparison_space": "encoded"})
    assert not o.passed and o.changed_pixels_outside == 1 and o.worst_pixel == (3, 3)
    assert o.locations == [(3, 3)] and abs(o.max_abs_delta_outside - 0.3) < 1e-6
    o = out({"before": img(4, 1, np.zeros(4)), "after": img(4, 1, [0.0, 0.2, 0.0, 0.5]),
             "allowed": msk(4, 1, np.zeros(4))}, {"comparison_space": "encoded"})
    assert o.changed_pixels_outside == 2 and o.worst_pixel == (3, 0)
    o = out({"before": img(2, 1, [0.0, 0.0]), "after": img(2, 1, [0.0, 0.05]),
             "allowed": msk(2, 1, [0.0, 0.0])},
            {"comparison_space": "encoded", "outside_threshold": 0.1})
    assert o.passed and abs(o.max_abs_delta_outside - 0.05) < 1e-6
    one = {"before": img(1, 1, [0.0]), "after": img(1, 1, [0.4]), "allowed": msk(1, 1, [0.3])}
    assert out(one, {"comparison_space": "encoded"}).passed
    o = out(one, {"comparison_space": "encoded", "coverage_epsilon": 0.5})
    assert not o.passed and o.changed_pixels_outside == 1
    srgb = {"before": img(1, 1, [0.4], "Srgb"), "after": img(1, 1, [0.5], "Srgb"),
            "allowed": msk(1, 1, [0.0])}
    enc = out(srgb, {"comparison_space": "encoded"}).max_abs_delta_outside
    lin = out(srgb, {"comparison_space": "decoded-linear"}).max_abs_delta_outside
    assert abs(enc - 0.1) < 1e-6 and lin < enc
    for sev in ("error", "warning", "metric"):       # tests.rs:307-337
        o = out({"before": img(4, 4, before), "after": img(4, 4, after),
                 "allowed": msk(4, 4, mk)}, {"comparison_space": "encoded", "severity": sev})
        assert not o.passed and o.severity == sev and o.fails_run() == (sev == "error")


def test_param_errors_through_the_abi(gpu):
    d = _dev(np.zeros((2, 2), np.float32))
    try:
        with pytest.raises(hip.HipCallError) as e:
            hip.assert_no_change_outside_mask_device(d.ptr, d.ptr, d.ptr, 2, 2, 1,
                                                     outside_threshold=-1.0)
        assert e.value.code == "E_ASSERT_PARAM" and e.value.error_class == "schema"
    finally:
        d.free()


def test_solver_candidate_leaves_outside_unchanged(gpu):
    """The blend's candidate against its target, scanned on the device at
    2048^2 RGBA: nothing outside the mask changes (the candidate starts as the
    target's bits), while the interior does."""
    from paintop_amd.device import DeviceArray
    from paintop_amd.inputs import seamless_clone_case

    n = 2048
    src, tgt, msk = seamless_clone_case(n, channels=4)
    out = DeviceArray(tgt.shape, np.float32)
    hip.poisson_solve_device(hip.make_desc(n, n, 4, 100000, 1e-6), src.ptr, tgt.ptr, msk.ptr,
                             out.ptr)
    r = hip.assert_no_change_outside_mask_device(tgt.ptr, out.ptr, msk.ptr, n, n, 4)
    assert r["changed_outside"] == 0 and r["max_outside"] == 0.0 and r["worst_pixel"] is None
    # the same scan with the mask inverted sees the interior change
    inv = DeviceArray.from_numpy((msk.numpy() <= 0.5).astype(np.float32))
    r2 = hip.assert_no_change_outside_mask_device(tgt.ptr, out.ptr, inv.ptr, n, n, 4)
    assert r2["changed_outside"] > 100000 and len(r2["locations"]) == 256
