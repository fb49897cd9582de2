"""CPU tests of assert.no_change_outside_mask@1: the oracle's scan against
the reference's own fixtures (crates/paintop-cpu/src/assert/tests.rs:148-306)
and the semantics scan_outside fixes (assert.rs:405-446), and the host op's
param and shape errors (raised before any device call)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from paintop_amd import ops
from paintop_amd.errors import ErrorClass, PaintopError
from paintop_amd.values import Extent, ImageDescriptor, MaskDescriptor, ResourceValue


def _img(w, h, c, samples, color="LinearSrgb"):
    lay = {1: "Gray", 3: "Rgb", 4: "Rgba"}[c]
    return ResourceValue(ImageDescriptor(Extent(w, h), lay, color), c, samples)


def _mask(w, h, samples):
    return ResourceValue(MaskDescriptor(Extent(w, h)), 1, samples)


def test_localized_edit_passes():  # tests.rs:148-171
    b = np.full((4, 4), 0.2, np.float32)
    a = b.copy()
    m = np.zeros((4, 4), np.float32)
    for i in (0, 1, 4, 5):
        a.flat[i] = 0.9
        m.flat[i] = 1.0
    r = O.outside_scan(b, a, m)
    assert r["changed_outside"] == 0 and r["max_outside"] == 0.0
    assert r["worst_pixel"] is None and r["locations"] == []


def test_one_pixel_leak_fails_with_worst_pixel():  # tests.rs:173-202
    b = np.full((4, 4), 0.2, np.float32)
    a = b.copy()
    m = np.zeros((4, 4), np.float32)
    for i in (0, 1, 4, 5):
        a.flat[i] = 0.9
        m.flat[i] = 1.0
    a.flat[15] = 0.5
    r = O.outside_scan(b, a, m)
    assert r["changed_outside"] == 1 and r["worst_pixel"] == (3, 3)
    assert r["locations"] == [(3, 3)] and abs(r["max_outside"] - 0.3) < 1e-6


def test_worst_pixel_is_the_largest_outside_delta():  # tests.rs:204-221
    r = O.outside_scan(np.zeros((1, 4), np.float32),
                       np.array([[0.0, 0.2, 0.0, 0.5]], np.float32), np.zeros((1, 4), np.float32))
    assert r["changed_outside"] == 2 and r["worst_pixel"] == (3, 0)
    assert r["locations"] == [(1, 0), (3, 0)]


def test_outside_threshold_tolerates_small_deltas():  # tests.rs:223-241
    r = O.outside_scan(np.zeros((1, 2), np.float32), np.array([[0.0, 0.05]], np.float32),
                       np.zeros((1, 2), np.float32), outside_threshold=0.1)
    assert r["changed_outside"] == 0 and abs(r["max_outside"] - 0.05) < 1e-6


def test_coverage_epsilon_widens_the_outside_region():  # tests.rs:243-273
    b, a, m = (np.array([[v]], np.float32) for v in (0.0, 0.4, 0.3))
    assert O.outside_scan(b, a, m)["changed_outside"] == 0
    assert O.outside_scan(b, a, m, coverage_epsilon=0.5)["changed_outside"] == 1


def test_decoded_linear_space_measures_a_light_delta():  # tests.rs:275-305
    b, a, m = (np.array([[v]], np.float32) for v in (0.4, 0.5, 0.0))
    enc = O.outside_scan(b, a, m)["max_outside"]
    lin = O.outside_scan(b, a, m, decode=True)["max_outside"]
    assert abs(enc - 0.1) < 1e-6 and lin < enc
    # IEC 61966-2-1 at the two samples (the decode the scan applies)
    dec = lambda c: ((c + 0.055) / 1.055) ** 2.4  # noqa: E731
    assert abs(lin - (dec(0.5) - dec(0.4))) < 1e-6


def test_scan_semantics_ties_nan_cap_and_channels():
    """First pixel wins a tie for the worst; NaN deltas never count; the
    location list keeps the first 256 leaks in row-major order while the count
    stays exact; a pixel's delta is its largest channel."""
    h, w, c = 30, 20, 3
    b = np.zeros((h, w, c), np.float32)
    a = b.copy()
    m = np.zeros((h, w), np.float32)
    a[..., 1] = 0.25                    # every pixel leaks by 0.25 in channel 1
    a[5, 7, 2] = 0.75                   # the largest delta, twice
    a[9, 3, 0] = -0.75
    a[0, 0, :] = np.nan                 # NaN: no delta at all
    m[2, :] = 1.0                       # row 2 is inside
    r = O.outside_scan(b, a, m, outside_threshold=0.1)
    n_out = h * w - w - 1
    assert r["changed_outside"] == n_out
    assert r["worst_pixel"] == (7, 5) and r["max_outside"] == 0.75
    locs = [(i % w, i // w) for i in range(h * w) if i // w != 2 and i != 0][:256]
    assert r["locations"] == locs
    # zero channels: the empty scan (assert.rs:407-409)
    z = O.outside_scan(np.zeros((2, 2, 0), np.float32), np.zeros((2, 2, 0), np.float32),
                       np.zeros((2, 2), np.float32))
    assert z["changed_outside"] == 0 and z["worst_pixel"] is None


@pytest.mark.parametrize("params,code", [
    ({"comparison_space": "perceptual"}, "E_ASSERT_PARAM"),   # tests.rs:393-406
    ({"severity": "fatal"}, "E_ASSERT_PARAM"),
    ({"outside_threshold": -1.0}, "E_ASSERT_PARAM"),
    ({"coverage_epsilon": math.inf}, "E_ASSERT_PARAM"),
    ({"coverage_epsilon": "0.1"}, "E_ASSERT_PARAM"),
])
def test_param_errors(params, code):
    op = ops.NoChangeOutsideMask()
    ins = {"before": _img(2, 2, 1, np.zeros(4)), "after": _img(2, 2, 1, np.zeros(4)),
           "allowed": _mask(2, 2, np.zeros(4))}
    with pytest.raises(PaintopError) as e:
        op.compute(ins, params)
    assert e.value.code == code and e.value.error_class == ErrorClass.SCHEMA


def test_shape_and_input_errors():  # tests.rs:378-391
    op = ops.NoChangeOutsideMask()
    ins = {"before": _img(2, 2, 1, np.zeros(4)), "after": _img(3, 2, 1, np.zeros(6)),
           "allowed": _mask(2, 2, np.zeros(4))}
    with pytest.raises(PaintopError) as e:
        op.compute(ins, {})
    assert e.value.code == "E_ASSERT_SHAPE" and e.value.error_class == ErrorClass.SEMANTIC
    ins["after"] = _img(2, 2, 1, np.zeros(4))
    ins["allowed"] = _mask(3, 2, np.zeros(6))
    with pytest.raises(PaintopError) as e:
        op.compute(ins, {})
    assert e.value.code == "E_ASSERT_SHAPE"
    del ins["allowed"]
    with pytest.raises(PaintopError) as e:
        op.compute(ins, {})
    assert e.value.code == "E_ASSERT_INPUT" and e.value.error_class == ErrorClass.REFERENCE


def test_mask_components_are_four_connected_interior_cells():
    """components.py: interior = mask > 0.5 off the image border
    (poisson.rs:340-355); diagonal contact does not join components; each box
    includes the one-cell Dirichlet ring."""
    from paintop_amd import components as C

    m = np.zeros((12, 12), np.float32)
    m[2:5, 2:5] = 1.0
    m[6:9, 6:9] = 1.0
    m[5, 5] = 1.0          # touches both squares only diagonally
    m[0, :] = 1.0          # the border row is never interior
    m[9:11, 1] = 0.5       # not > 0.5
    labels, comps = C.components(m)
    assert [(c.y0, c.y1, c.x0, c.x1, c.cells) for c in comps] == [
        (1, 6, 1, 6, 9), (4, 7, 4, 7, 1), (5, 10, 5, 10, 9)]
    assert labels[0].max() == 0 and labels[5, 5] == 2
    assert (labels > 0).sum() == 19
