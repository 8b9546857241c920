"""CPU checks of the learned linear start's basis, predictor and fit.

The real-DFT basis is pinned to numpy's FFT (cosine rows are the real part,
sine rows minus the imaginary part, DC and Nyquist at 1/sqrt(16)), the fit to
responses built in pixel space without the basis (a uniform gain on the
centred data recovers A everywhere off DC, a gain on gamma recovers B, a
constant shift of k sigma recovers c = k)."""
import os

import numpy
import pytest

from ncs_amd import linstart


def _tiles(n, seed):
    rng = numpy.random.default_rng(seed)
    d = rng.gamma(2.0, 40.0, (n, 16, 16))
    g = rng.uniform(0.5, 8.0, (n, 16, 16))
    return d, g


def test_basis_is_orthonormal_and_matches_the_fft():
    R = linstart.real_dft_basis()
    assert numpy.allclose(R @ R.T, numpy.eye(16), atol=1e-13)
    x = numpy.random.default_rng(1).standard_normal(16)
    F = numpy.fft.fft(x)
    y = R @ x
    assert numpy.isclose(y[0], F[0].real / 4.0)
    assert numpy.isclose(y[8], F[8].real / 4.0)
    for k in range(1, 8):
        assert numpy.isclose(y[k], numpy.sqrt(2.0 / 16) * F[k].real)
        assert numpy.isclose(y[8 + k], -numpy.sqrt(2.0 / 16) * F[k].imag)


def test_spec_and_pix_are_inverse_and_keep_energy():
    x = numpy.random.default_rng(2).standard_normal((5, 16, 16))
    z = linstart.spec(x)
    assert numpy.allclose(linstart.pix(z), x, atol=1e-12)
    assert numpy.allclose((z ** 2).sum(axis=(1, 2)), (x ** 2).sum(axis=(1, 2)))
    # the 2-D transform is the 1-D one along both axes
    one = linstart.spec(numpy.ones((16, 16)))
    assert numpy.isclose(one[0, 0], 16.0) and numpy.abs(one).sum() == pytest.approx(16.0)


def test_fit_recovers_gains_built_in_pixel_space():
    d, g = _tiles(3000, 3)
    mu = d.mean(axis=(1, 2), keepdims=True)
    sig = numpy.sqrt(numpy.maximum(mu, 1.0) + g.mean(axis=(1, 2), keepdims=True))
    u = d + 0.3 * (d - mu) + 0.5 * g / sig + 2.0 * sig
    coef = linstart.fit(d, g, u).astype(numpy.float64)
    A = coef[:256].reshape(16, 16)
    B = coef[256:512].reshape(16, 16)
    off_dc = numpy.ones((16, 16), bool)
    off_dc[0, 0] = False
    assert numpy.allclose(A[off_dc], 0.3, atol=1e-5)
    assert numpy.allclose(B, 0.5, atol=1e-5)
    assert coef[512] == pytest.approx(2.0, abs=1e-5)
    # the predictor reproduces the response from the fitted coefficients
    assert numpy.allclose(linstart.predict(d, g, coef), u, rtol=1e-5, atol=1e-3)


def test_fit_separates_per_coefficient():
    """A response in one spectral coefficient only leaves every other
    coefficient's gains at zero."""
    d, g = _tiles(2000, 4)
    X0, X1, sig = linstart.features(d, g)
    Z = numpy.zeros_like(X0)
    Z[:, 3, 11] = 0.7 * X0[:, 3, 11] - 0.2 * X1[:, 3, 11]
    u = d + sig * linstart.pix(Z)
    coef = linstart.fit(d, g, u).astype(numpy.float64)
    A = coef[:256].reshape(16, 16)
    B = coef[256:512].reshape(16, 16)
    assert A[3, 11] == pytest.approx(0.7, abs=1e-5) and B[3, 11] == pytest.approx(-0.2, abs=1e-5)
    A[3, 11] = B[3, 11] = 0.0
    assert numpy.abs(A).max() < 1e-5 and numpy.abs(B).max() < 1e-5 and abs(coef[512]) < 1e-5


def test_stored_coefficients_load():
    if not os.path.exists(linstart.DEFAULT_COEF):
        pytest.skip("no stored coefficients")
    c = linstart.load()
    assert c.shape == (linstart.N_COEF,) and c.dtype == numpy.float32
    assert numpy.all(numpy.isfinite(c))


def test_fit_slope_recovers_slopes_built_in_pixel_space():
    """Solutions whose gains move with ln(alpha / 0.1): a uniform gain
    a(alpha) = 0.3 + 0.05 ln(alpha / 0.1) on the centred data, b(alpha) = 0.5
    - 0.1 ln(alpha / 0.1) on gamma and a shift of (2 + 0.4 ln(alpha / 0.1))
    sigma.  Given the alpha = 0.1 fit, fit_slope recovers the three slopes
    (A' everywhere off DC, B', c')."""
    d, g = _tiles(2000, 5)
    mu = d.mean(axis=(1, 2), keepdims=True)
    sig = numpy.sqrt(numpy.maximum(mu, 1.0) + g.mean(axis=(1, 2), keepdims=True))

    def solution(alpha):
        ell = numpy.log(alpha / 0.1)
        return (d + (0.3 + 0.05 * ell) * (d - mu) + (0.5 - 0.1 * ell) * g / sig
                + (2.0 + 0.4 * ell) * sig)

    coef = linstart.fit(d, g, solution(0.1))
    slope = linstart.fit_slope(d, g, [(a, solution(a)) for a in (0.05, 0.2, 0.3)],
                               coef).astype(numpy.float64)
    A1 = slope[:256].reshape(16, 16)
    B1 = slope[256:512].reshape(16, 16)
    off_dc = numpy.ones((16, 16), bool)
    off_dc[0, 0] = False
    assert numpy.allclose(A1[off_dc], 0.05, atol=1e-5)
    assert numpy.allclose(B1, -0.1, atol=1e-5)
    assert slope[512] == pytest.approx(0.4, abs=1e-5)
    # the predictor at each alpha then reproduces the solutions
    for a in (0.05, 0.3):
        u = linstart.predict(d, g, coef, alpha=a, slope=slope)
        assert numpy.allclose(u, solution(a), rtol=1e-6, atol=1e-4)


def test_predictor_at_the_reference_alpha_ignores_the_slopes():
    d, g = _tiles(50, 6)
    rng = numpy.random.default_rng(7)
    coef = rng.uniform(-0.3, 0.3, linstart.N_COEF)
    slope = rng.uniform(-0.1, 0.1, linstart.N_COEF)
    u0 = linstart.predict(d, g, coef)
    assert numpy.array_equal(linstart.predict(d, g, coef, alpha=linstart.ALPHA_REF,
                                              slope=slope), u0)
    assert not numpy.allclose(linstart.predict(d, g, coef, alpha=0.2, slope=slope), u0)
    # alpha <= 0 keeps the coefficients; beyond 8x alpha_ref the slope is clamped
    assert numpy.array_equal(linstart.predict(d, g, coef, alpha=0.0, slope=slope), u0)
    assert numpy.array_equal(linstart.predict(d, g, coef, alpha=5.0, slope=slope),
                             linstart.predict(d, g, coef, alpha=0.8, slope=slope))


def test_stored_slopes_load():
    s = linstart.load_slope()
    if s is None:
        pytest.skip("no stored slopes")
    assert s.shape == (linstart.N_COEF,) and s.dtype == numpy.float32
    assert numpy.all(numpy.isfinite(s))
