"""Pins the oracle's masked-likelihood solve (ncso_solve_tiles_masked_f32),
the checker of the zero-padded block tiling, before it is trusted.

A padded pixel has likelihood weight lin = 0: it stays a variable of the
noise term only.  Each check comes from outside the oracle and fails a
plausible masking mistake (a padded pixel kept in the likelihood value or
its gradient, its data read anyway, the weight applied to the wrong pixel):
  * lin = 1 everywhere reproduces the unmasked solve bit for bit;
  * an all-ones mask makes every pixel separable (Parseval): valid pixels
    reach the per-pixel closed form, padded pixels go to 0 whatever data
    they hold;
  * a DC term on top of a uniform mask couples every pixel, padded ones
    included, with a positive-definite Hessian: the minimiser of Newton's
    method in FP64.
A pure-DC mask is not used: with padded pixels free of the likelihood its
minimum is not unique (only their sum is fixed), and a first step along
-(1, ..., 1) takes valid pixels into the 1e-6 guard, where the reference's
gradient is not the cost's derivative, so L-BFGS (scipy's L-BFGS-B too)
stops early there.
"""
import numpy
import pytest

import oracle
from ncs_amd import masks


def _tiles(seed, n, pad_rows, pad_cols, pad_value=5.0):
    """Tiles whose last pad_rows rows / pad_cols cols are padding (as an edge
    block of a frame); padded pixels hold pad_value data and gamma, which a
    correct solve must ignore."""
    rng = numpy.random.default_rng(seed)
    d = rng.uniform(2.0, 60.0, (n, 16, 16)).astype(numpy.float32)
    g = rng.uniform(1.0, 8.0, (n, 16, 16)).astype(numpy.float32)
    lin = numpy.ones((n, 16, 16), numpy.float32)
    for t in range(n):
        pr, pc = pad_rows[t % len(pad_rows)], pad_cols[t % len(pad_cols)]
        if pr:
            lin[t, 16 - pr:, :] = 0.0
        if pc:
            lin[t, :, 16 - pc:] = 0.0
    d[lin == 0] = pad_value
    g[lin == 0] = pad_value
    return d, g, lin


@pytest.mark.parametrize("kind", ["otf", "random"])
def test_all_valid_pixels_reproduce_the_unmasked_solve(kind):
    rng = numpy.random.default_rng(3)
    d = rng.uniform(2.0, 60.0, (12, 16, 16)).astype(numpy.float32)
    g = rng.uniform(1.0, 8.0, (12, 16, 16)).astype(numpy.float32)
    mask = masks.otf_mask() if kind == "otf" else masks.random_mask(16, rng)
    mf = masks.to_fft_order(mask).astype(numpy.float32)
    a = oracle.solve_tiles_masked(d, g, numpy.ones_like(d), mf, 0.1)
    b = oracle.solve_tiles(d, g, mf, 0.1, f32=True)
    assert numpy.array_equal(a["u"], b["u"].astype(numpy.float32))
    assert numpy.array_equal(a["cost"], b["cost"])
    assert numpy.array_equal(a["status"], b["status"])
    assert numpy.array_equal(a["iters"], b["iters"])


def test_all_ones_mask_closed_form_on_valid_pixels_zero_on_padding():
    """NC = sum u^2 with an all-ones mask, so each pixel solves
    2 a u^2 + (1 + 2 a gamma) u - d = 0 alone; a padded pixel has only
    a u^2 and goes to 0."""
    alpha = 0.3
    d, g, lin = _tiles(5, 8, (0, 3, 9, 15), (0, 1, 7, 14))
    p = oracle.default_param(epsilon=1e-7)
    res = oracle.solve_tiles_masked(d, g, lin, numpy.ones((16, 16), numpy.float32), alpha,
                                    param=p)
    dd, gg = d.astype(numpy.float64), g.astype(numpy.float64)
    b = 1.0 + 2.0 * alpha * gg
    closed = (-b + numpy.sqrt(b * b + 8.0 * alpha * dd)) / (4.0 * alpha)
    expect = numpy.where(lin > 0, closed, 0.0)
    assert numpy.all(res["status"] == 0)
    err = numpy.max(numpy.abs(res["u"] - expect)[lin > 0])
    assert err < 1e-5, err  # float32 output of values up to ~40
    assert numpy.max(numpy.abs(res["u"][lin == 0])) < 1e-6  # from 5.0


def _newton_masked(d, g, lin, b2, m2, alpha):
    """FP64 Newton with backtracking on
    f(u) = sum_lin [u - c ln(u + g)] + alpha (b2 sum u^2 + m2 (sum u)^2 / 256)."""
    c = (d + g) * lin
    x = numpy.where(lin > 0, d, 0.0)

    def f(x):
        v = numpy.where(lin > 0, x + g, 1.0)
        return (numpy.sum(lin * x - c * numpy.log(v))
                + alpha * (b2 * numpy.sum(x * x) + m2 * numpy.sum(x) ** 2 / 256.0))

    for _ in range(60):
        v = numpy.where(lin > 0, x + g, 1.0)
        grad = lin - c / v + 2 * alpha * (b2 * x + m2 * numpy.sum(x) / 256.0)
        H = numpy.diag(c / v ** 2 + 2 * alpha * b2) + 2 * alpha * m2 / 256.0
        step = numpy.linalg.solve(H, grad)
        t = 1.0
        while numpy.any((x - t * step + g)[lin > 0] <= 0) or f(x - t * step) > f(x):
            t *= 0.5
            if t < 1e-12:
                break
        x = x - t * step
        if numpy.max(numpy.abs(t * step)) < 1e-13 * max(1.0, numpy.max(numpy.abs(x))):
            break
    return x, f(x)


def test_dc_coupling_matches_newton_fp64():
    """Mask b everywhere plus a DC term: NC = b^2 sum u^2 + m^2 (sum u)^2 / N^2
    couples every pixel, padded ones included, with a positive-definite
    Hessian; the oracle's minimiser against Newton's in FP64."""
    alpha, b, m = 0.1, 0.4, 0.9
    d, g, lin = _tiles(21, 6, (0, 4, 11), (2, 0, 13))
    mf = numpy.full((16, 16), b, numpy.float32)
    mf[0, 0] = numpy.float32(numpy.sqrt(b * b + m * m))
    b2 = float(numpy.float32(b)) ** 2
    m2 = float(mf[0, 0]) ** 2 - b2
    res = oracle.solve_tiles_masked(d, g, lin, mf, alpha,
                                    param=oracle.default_param(epsilon=1e-10))
    for t in range(len(d)):
        x, fx = _newton_masked(d[t].ravel().astype(numpy.float64),
                               g[t].ravel().astype(numpy.float64),
                               lin[t].ravel().astype(numpy.float64), b2, m2, alpha)
        err = numpy.max(numpy.abs(res["u"][t].ravel() - x))
        assert err < 1e-4, (t, err)
        assert abs(res["cost"][t] - fx) <= 1e-9 * abs(fx), (t, res["cost"][t], fx)
