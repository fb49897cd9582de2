"""The CPU reference U-Net (oracle/unet_ref.py) is pinned against central
finite differences of its own forward pass, and its parameter layout against
the library's (the C-ABI library loads without a GPU)."""
import numpy as np

from oracle import unet_ref as U


def _fd(p, x, y, w, v, eps):
    lp, _, _ = U.loss_and_grads(p + eps * v, x, y, w)
    lm, _, _ = U.loss_and_grads(p - eps * v, x, y, w)
    return (lp - lm) / (2 * eps)


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    p = U.init_params(1).astype(np.float64)
    B, R = 1, 16
    x = rng.standard_normal((B, R, R, 4))
    x[..., 3] = 0
    y = rng.standard_normal((B, R, R))
    w = rng.random((B, R, R))
    loss, g, _ = U.loss_and_grads(p, x, y, w)
    lay, hoff, total = U.layout()
    for li in range(len(U.LAYERS)):
        ow, ob, cout, cp, kp = lay[li]
        v = np.zeros(total)
        blk = np.zeros((cout, kp))
        r = rng.standard_normal((cout, 9, cp))
        blk[:, :9 * cp] = r.reshape(cout, -1)
        v[ow:ow + cout * kp] = blk.ravel()
        v[ob:ob + cout] = rng.standard_normal(cout)
        an = float(g @ v)
        # a ReLU kink inside the step spoils one step size, not both
        errs = [abs(_fd(p, x, y, w, v, e) - an) / max(abs(an), 1e-12) for e in (1e-7, 3e-8)]
        assert min(errs) < 1e-5, (li, errs)
    v = np.zeros(total)
    v[hoff:total] = rng.standard_normal(total - hoff)
    assert abs(_fd(p, x, y, w, v, 1e-7) - g @ v) < 1e-6 * abs(g @ v)


def test_reference_padding_gradients_are_zero():
    rng = np.random.default_rng(4)
    p = U.init_params(2)
    x = rng.standard_normal((1, 16, 16, 4))
    x[..., 3] = 0
    _, g, _ = U.loss_and_grads(p, x, rng.standard_normal((1, 16, 16)), np.ones((1, 16, 16)))
    lay, _, _ = U.layout()
    ow, ob, cout, cp, kp = lay[0]
    g0 = g[ow:ow + cout * kp].reshape(cout, kp)
    assert kp > 9 * cp and np.all(g0[:, 9 * cp:] == 0)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-39],
                    dtype=np.float32)
    r = U.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0  # halfway between 1 and 1 + 2^-7: even (1.0)
    assert r[2] == 1.0 + 2 ** -6  # halfway: rounds to the even neighbour 1 + 2^-6
    assert r[3] == 1.0
    assert r[4] == -2.5


def test_layout_matches_library():
    from paintop_amd import hip

    _, _, total = U.layout()
    assert hip.unet_param_storage() == total
    assert hip.unet_param_count() == U.param_count() == 302993


def test_reference_pyramid_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(12)
    p = U.init_params(4).astype(np.float64)
    x = rng.standard_normal((2, 32, 32, 4))
    y = rng.standard_normal((2, 32, 32))
    w = rng.random((2, 32, 32))
    _, g, _ = U.loss_and_grads(p, x, y, w, pyramid=0.7)
    _, hoff, total = U.layout()
    v = np.zeros(total)
    v[hoff:total] = rng.standard_normal(total - hoff)  # head: smooth in the loss
    lp, _, _ = U.loss_and_grads(p + 1e-6 * v, x, y, w, pyramid=0.7)
    lm, _, _ = U.loss_and_grads(p - 1e-6 * v, x, y, w, pyramid=0.7)
    assert abs((lp - lm) / 2e-6 - g @ v) < 1e-6 * abs(g @ v)
    # the multi-scale term of a constant error: each level adds the same amount
    l0, _, out = U.loss_and_grads(p, x, y, np.ones_like(w))
    l1, _, _ = U.loss_and_grads(p, x, out - 0.25, np.ones_like(w), pyramid=1.0)
    assert abs(l1 - 5 * 0.25 ** 2) < 1e-12
