"""The stop rules of poisson::solve (crates/paintop-cpu/src/poisson.rs:168-240)
on the GPU, for the paths the ordinary convergence tests never reach:

* Stalled: a tolerance below the FP64 floor of the residual (1e-300 is inside
  the contract's (0, 1)) ends every solve as stalled once the residual stops
  improving. With a caller-fixed omega the window is the reference's 16 sweeps
  (poisson.rs:81); with the tuned omega it is max(16, ceil(3 / (2 - omega))),
  because red/black SOR near omega = 2 can go ~1/(2 - omega) sweeps without a
  new best while it is still converging. The reported history is replayed
  through the reference's rule (no new best below best * (1 - 1e-12),
  poisson.rs:215-228) with that window: the rule must not have fired before
  the last sweep, and the solve must have reached the FP64 floor first.
* Non-finite input: a NaN in the guidance inside the mask makes r0 NaN; the
  reference then runs 16 sweeps whose residuals are all NaN and stops as
  stalled (max-iterations when the cap is below 16), and the op reports
  all_finite = false (poisson_blend.rs:573).
* History entries: between FP64 refinements the per-sweep entries are the
  solver's FP32 estimates (the residual identity of the fused sweep). Each is
  compared with the exact FP64 residual of the same sweep, which a run capped
  at that sweep reports as its final residual (the capped run performs the
  same sweeps and refinements up to the cap).
"""
import math

import numpy as np
import pytest

from paintop_amd import hip, ops
from paintop_amd.inputs import gaussian_bump_case
from paintop_amd.values import image, mask
from tests.helpers import classify

pytestmark = pytest.mark.gpu

STALL_IMPROVEMENT = 1e-12  # poisson.rs:85


def replay_stall(history, window):
    """The reference's stall guard over a history: the 1-based sweep at which
    it fires, or None."""
    best, since = math.inf, 0
    for k, rel in enumerate(history, start=1):
        if rel < best * (1.0 - STALL_IMPROVEMENT):
            best, since = rel, 0
        else:
            since += 1
            if since >= window:
                return k
    return None


def _bump(n):
    src, tgt, m = gaussian_bump_case(n)
    return src, tgt, m


@pytest.mark.parametrize("tuned", [False, True])
def test_stall_below_the_fp64_floor(gpu, tuned):
    src, tgt, m = _bump(128)
    n = m.shape[0]
    omega = 0.0 if tuned else 1.8  # the reference's fixed OMEGA (poisson_blend.rs:78)
    desc = hip.make_desc(n, n, 1, 50000, 1e-300, omega=omega)
    out, r = gpu.poisson_solve(desc, src, tgt, m)
    window = max(16, math.ceil(3.0 / (2.0 - r.omega))) if tuned else 16
    print(f"[stop] tuned={tuned}: omega {r.omega:.6f}, window {window}, {r.iterations} sweeps, "
          f"{r.stop_reason}, final {r.final_residual:.3e}, refinements {r.refinements}",
          flush=True)
    assert r.stop_reason == "stalled" and not r.converged
    assert r.iterations < 50000 and len(r.residual_history) == r.iterations
    # it reached the FP64 floor (a relative residual of ~1e-13..1e-14 here)
    # before the guard fired, and the final residual is the FP64 value
    assert r.final_residual < 1e-11
    assert r.residual_history[-1] == r.final_residual
    if tuned:
        assert window > 16  # the case exercises the tuned window
    h = r.residual_history
    # the rule did not fire before the last sweep, and it is one sweep (the
    # FP64-confirmed entry) from firing at the end
    assert replay_stall(h[:-1], window) is None
    assert replay_stall(h, window) in (None, len(h))
    best = min(h[:-window]) if len(h) > window else math.inf
    assert all(v >= best * (1.0 - STALL_IMPROVEMENT) for v in h[-window:-1])
    inside = classify(m).astype(bool)
    assert out.reshape(n, n)[~inside].tobytes() == tgt[~inside].tobytes()


@pytest.mark.parametrize("cap,iters,reason", [(100, 16, "stalled"), (16, 16, "stalled"),
                                               (5, 5, "max-iterations")])
def test_nan_guidance_inside_the_mask(gpu, cap, iters, reason):
    src, tgt, m = _bump(96)
    n = m.shape[0]
    src = src.copy()
    src[n // 2, n // 2] = np.nan  # inside the ellipse: rhs, and so r0, are NaN
    op = ops.PoissonBlend()
    out = op.compute({"source": image(n, n, src), "target": image(n, n, tgt),
                      "mask": mask(n, n, m)}, {"max_iterations": cap, "tolerance": 1e-6})
    rep = out["report"].as_report()
    sd = rep.solver
    cand = out["candidate"].array()[..., 0]
    print(f"[stop] NaN guidance, cap {cap}: {sd.iterations} sweeps, {sd.stop_reason}, "
          f"all_finite {rep.all_finite}", flush=True)
    assert sd.iterations == iters and sd.stop_reason == reason and sd.converged is False
    assert len(sd.residual_history) == iters and all(math.isnan(v) for v in sd.residual_history)
    assert math.isnan(sd.final_residual)
    assert rep.all_finite is False
    inside = classify(m).astype(bool)
    assert cand[~inside].tobytes() == tgt[~inside].tobytes()
    assert np.isnan(cand[inside]).any()


def test_nan_outside_the_stencil_keeps_solving(gpu):
    """A NaN in the target far from the mask is neither an unknown nor
    Dirichlet data of one: the solve converges and the NaN passes through."""
    src, tgt, m = _bump(96)
    n = m.shape[0]
    tgt = tgt.copy()
    tgt[2, 2] = np.nan
    op = ops.PoissonBlend()
    out = op.compute({"source": image(n, n, src), "target": image(n, n, tgt),
                      "mask": mask(n, n, m)}, {"max_iterations": 5000, "tolerance": 1e-6})
    rep = out["report"].as_report()
    cand = out["candidate"].array()[..., 0]
    assert rep.solver.converged and rep.all_finite is False
    inside = classify(m).astype(bool)
    assert cand[~inside].tobytes() == tgt[~inside].tobytes()
    assert np.isfinite(cand[inside]).all()


def test_history_estimates_match_the_fp64_residual(gpu):
    """history[m-1] of a long run (an FP32 estimate unless sweep m was a
    refinement point) against the FP64 residual of the same sweep, which a run
    capped at m reports as final_residual. Bound: 1e-3 relative, the accuracy
    of an FP32 evaluation of the correction residual while the inner residual
    is within the 1e3-fold refinement window."""
    src, tgt, m = _bump(192)
    n = m.shape[0]
    long_desc = hip.make_desc(n, n, 1, 20000, 1e-10)
    _, rl = gpu.poisson_solve(long_desc, src, tgt, m)
    h = rl.residual_history
    assert rl.converged and len(h) > 200
    worst = 0.0
    for mm in (3, 17, 50, 123, 200, len(h) // 2, len(h) - 5):
        _, rc = gpu.poisson_solve(hip.make_desc(n, n, 1, mm, 1e-10), src, tgt, m)
        assert rc.iterations == mm and rc.stop_reason == "max-iterations"
        assert rc.residual_history[:-1] == h[:mm - 1]  # the same trajectory up to m
        d = abs(h[mm - 1] - rc.final_residual) / rc.final_residual
        worst = max(worst, d)
        print(f"[stop] sweep {mm}: estimate {h[mm - 1]:.6e} fp64 {rc.final_residual:.6e} "
              f"rel diff {d:.2e}", flush=True)
        assert d <= 1e-3
    print(f"[stop] worst relative difference of an estimate: {worst:.2e}", flush=True)
