"""Pins of the oracle's large-size reference solve (oracle/tight_solve.c) and
of the CPU generators the full-size fixtures are rebuilt with.

The tight solve is checked against what the mathematics fixes independently
of any iteration: a sparse direct LU solve of the same 5-point system
(tests/helpers.direct_solve) on non-square random domains for several lambda,
an exactly representable harmonic solution (a linear field), the zero-residual
start, and bit-identical results for 1 and several threads. A wrong sign, a
dropped lambda term or a transposed coupling in either the operator or the
multigrid preconditioner's Galerkin coarsening fails the direct-solve
comparison, or the closed form.
"""
import math
import zlib

import numpy as np
import pytest

from tests.helpers import classify, direct_solve, laplacian, residual_norm


def _random_domain(h, w, seed):
    rng = np.random.default_rng(seed)
    m = (rng.random((h, w)) < 0.8).astype(np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    m *= (((xx - w / 2) / (0.45 * w)) ** 2 + ((yy - h / 2) / (0.4 * h)) ** 2 <= 1).astype(np.float32)
    return m


@pytest.mark.parametrize("lam", [0.0, 0.01, 0.37, 5.0])
@pytest.mark.parametrize("shape", [(37, 53), (64, 48), (90, 121)])
def test_tight_solve_matches_direct_solve(oracle, lam, shape):
    h, w = shape
    rng = np.random.default_rng(h * w)
    m = _random_domain(h, w, h + w)
    cells = classify(m)
    src = rng.random((h, w)).astype(np.float32)
    tgt = rng.random((h, w)).astype(np.float32)
    rhs = np.where(cells == 1, laplacian(src), 0.0)
    u, rep = oracle.tight_solve(cells, rhs, tgt.astype(np.float64), lam, tolerance=1e-12)
    d = direct_solve(cells, rhs, tgt.astype(np.float64), lam)
    assert rep["final_residual"] <= 1e-12, rep
    assert np.abs(u - d).max() < 1e-10
    # outside the interior the field is the anchor, bit for bit
    assert u[cells == 0].tobytes() == tgt.astype(np.float64)[cells == 0].tobytes()
    # the reported residual is the independent numpy residual of the field
    r0 = residual_norm(cells, rhs, tgt, lam, tgt)
    assert rep["r0"] == pytest.approx(r0, rel=1e-12)
    assert residual_norm(cells, rhs, tgt, lam, u) / r0 <= 2e-12


def test_tight_solve_linear_field_is_exact(oracle):
    # a linear field is harmonic: with rhs = 0 and the boundary taken from
    # 2x + 3y the solution is 2x + 3y inside as well (poisson/tests.rs:54-116)
    h, w = 70, 90
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    lin = 2.0 * xx + 3.0 * yy
    m = _random_domain(h, w, 3)
    cells = classify(m)
    anchor = np.where(cells == 1, 0.25, lin)  # interior starts away from the answer
    u, rep = oracle.tight_solve(cells, np.zeros((h, w)), anchor, 0.0, tolerance=1e-13)
    assert np.abs(u - lin).max() < 1e-9


def test_tight_solve_zero_residual_start(oracle):
    h, w = 40, 40
    cells = classify(np.ones((h, w), np.float32))
    a = np.full((h, w), 0.5)
    u, rep = oracle.tight_solve(cells, np.zeros((h, w)), a, 0.0)
    assert rep["r0"] == 0.0 and rep["iterations"] == 0 and rep["converged"]
    assert u.tobytes() == a.tobytes()


def test_tight_solve_bit_identical_across_threads(oracle):
    h, w = 96, 80
    rng = np.random.default_rng(5)
    cells = classify(_random_domain(h, w, 9))
    rhs = np.where(cells == 1, rng.standard_normal((h, w)) * 1e-3, 0.0)
    anchor = rng.random((h, w))
    u1, r1 = oracle.tight_solve(cells, rhs, anchor, 0.0, threads=1)
    u4, r4 = oracle.tight_solve(cells, rhs, anchor, 0.0, threads=4)
    assert u1.tobytes() == u4.tobytes() and r1["iterations"] == r4["iterations"]


def _fma(a, b, c):
    """a*b + c rounded once (exact rational arithmetic, then one rounding)."""
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def test_polygon_mask_matches_host_ray_cast(oracle):
    from paintop_amd.inputs import star_polygon

    w, h = 70, 50
    verts = star_polygon(35, 25, 22, 0.5, 6, 0.1, 5)
    got = oracle.polygon_mask(w, h, verts)

    def inside(x, y):  # the half-open ray cast of mask_polygon.rs:161-195, nonzero rule
        wn = 0
        for i in range(len(verts)):
            (x0, y0), (x1, y1) = verts[i], verts[(i + 1) % len(verts)]
            b0, b1 = y0 <= y, y1 <= y
            if b0 == b1:
                continue
            t = (y - y0) / (y1 - y0)
            if _fma(t, x1 - x0, x0) > x:
                wn += 1 if b0 else -1
        return wn != 0

    want = np.array([[1.0 if inside(x + 0.5, y + 0.5) else 0.0 for x in range(w)]
                     for y in range(h)], np.float32)
    assert got.tobytes() == want.tobytes()
    # a square (the screened workload's border mask) is its exact pixel box
    sq = oracle.polygon_mask(20, 20, [(4, 4), (16, 4), (16, 16), (4, 16)])
    assert sq.sum() == 144 and sq[4:16, 4:16].all()


def test_gauss_noise_formula_and_statistics(oracle):
    from paintop_amd.inputs import mix64

    n = 200_000
    out = np.zeros(n, np.float32)
    oracle.gauss_noise(out, 403, 0.01)
    G = 0x9E3779B97F4A7C15
    for i in (0, 1, 17, 12345, n - 1):  # the formula, restated with the host libm
        a = mix64((403 + G * (2 * i + 1)) & 0xFFFFFFFFFFFFFFFF)
        b = mix64((403 + G * (2 * i + 2)) & 0xFFFFFFFFFFFFFFFF)
        u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740992.0)
        u2 = (b >> 11) * (1.0 / 9007199254740992.0)
        z = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        assert out[i] == np.float32(0.01 * z)
    assert abs(out.mean()) < 1e-4 and abs(out.std() - 0.01) < 1e-4


@pytest.mark.parametrize("name", ["seamless_clone_4096_tight.npz", "screened_8192_tight.npz"])
def test_full_size_fixtures_are_consistent(name):
    """The committed fixtures: every recorded solve reached 1e-12 (or the FP64
    floor just above it), samples lie on interior cells of the recorded
    classification size, and block means average the sampled field sensibly."""
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", name)
    z = np.load(path)
    n = int(z["size"])
    pts = z["pts"]
    assert pts.size == 4096 and pts.min() >= 0 and pts.max() < n * n
    chans = sorted(k[:-6] for k in z.files if k.endswith("_u_pts"))
    assert chans
    for c in chans:
        r0, fin, its, _ = z[f"{c}_report"]
        assert r0 > 0 and fin <= 2e-12 and its > 0
        u = z[f"{c}_u_pts"]
        assert np.isfinite(u).all()
        # harmonic-type blend of [0, 1] data stays near [0, 1]
        assert u.min() > -0.5 and u.max() < 1.5
        blk_n = z[f"{c}_blk_n"]
        assert int(blk_n.sum()) == int(z["n_interior"])


@pytest.mark.parametrize("n", [96, 160, 256])
def test_sor_projection_brackets_the_lexicographic_solve(oracle, n):
    from tests.helpers import smallest_eigenvalue, star_case

    """The projections the benchmark reports for cpu.reference at the full
    size, which it cannot run to tolerance, checked on the same workload
    scaled down, where the restated solve (or_solve, omega = 1.8) runs to 1e-6:
    Young's spectral radius of lexicographic SOR from the domain's smallest
    eigenvalue (scipy eigsh of the 5-point operator), anchored at the measured
    residual after 40 sweeps, over-estimates the sweep count by at most 25%
    here (every residual component decays at least as fast as the spectral
    radius); the count projected from a relative residual of 1 is 1.2-2x the
    actual one. mu is recovered exactly from the optimal omega."""
    cells, rhs, t = star_case(n)
    mu = smallest_eigenvalue(cells)
    _, rep = oracle.solve(cells, rhs, t, 0.0, 100000, 1e-6, 1.8)
    assert rep.converged
    h = rep.residual_history
    rho = oracle.sor_spectral_radius(mu, 0.0, 1.8)
    anchored = 40 + math.log(1e-6 / h[39]) / math.log(rho)
    from_one = oracle.project_sor_sweeps(mu, 0.0, 1.8, 1e-6)
    print(f"n={n}: mu {mu:.4e}, solve {rep.iterations}, anchored {anchored:.0f}, "
          f"from 1 {from_one:.0f}")
    assert rep.iterations <= anchored <= 1.25 * rep.iterations
    assert 1.2 * rep.iterations <= from_one <= 2.0 * rep.iterations
    w_opt = 2.0 / (1.0 + math.sqrt(1.0 - ((4.0 - mu) / 4.0) ** 2))
    assert oracle.mu_from_optimal_omega(w_opt) == pytest.approx(mu, rel=1e-6)
