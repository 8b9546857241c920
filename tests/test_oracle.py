"""Pins the FP64 oracle (oracle/) before it is trusted as the checker.

Every check comes from outside the oracle: numpy.fft (a library DFT),
central finite differences, closed forms (alpha = 0, Parseval with an
all-ones mask, constant tiles under a DC-free mask, a pure-DC mask solved by
Newton's method), scipy's L-BFGS-B, the golden fixtures of
tests/golden/make_golden.py, and the reference's own two-tilings test
(clib/pyCNCS/test/test_im.py:12-38).
"""
import numpy
import pytest
import scipy.optimize

import oracle


def _np_cost(u, d, g, mask_fft, alpha):
    v = numpy.maximum(u + g, 1e-6)
    U = numpy.fft.fft2(u)
    return (numpy.sum(u - (d + g) * numpy.log(v))
            + alpha * numpy.sum(mask_fft ** 2 * numpy.abs(U) ** 2) / u.size)


@pytest.mark.parametrize("mask", ["otf", "gauss", "random"])
def test_cost_and_gradient_match_golden(golden, mask):
    mf = golden["mask_" + mask]
    for i in range(golden["eval_u"].shape[0]):
        f, g = oracle.cost_grad(golden["eval_u"][i], golden["eval_d"][i],
                                golden["eval_g"][i], mf, 0.1)
        fr = golden["eval_cost_" + mask][i]
        assert abs(f - fr) <= 1e-11 * abs(fr)
        gr = golden["eval_grad_" + mask][i]
        assert numpy.max(numpy.abs(g - gr)) <= 1e-9 * max(1.0, numpy.max(numpy.abs(gr)))


def test_pycncs_subregion_api_matches_numpy():
    """clib/pyCNCS/test/test_sr.py:12-53 with numpy as the other side."""
    rng = numpy.random.default_rng(3)
    sr = oracle.NCSCSubRegion(16)
    for _ in range(5):
        gamma = rng.uniform(2, 4, (16, 16))
        image = rng.uniform(0.01, 10, (16, 16))
        u = rng.uniform(0.01, 10, (16, 16))
        m = oracle_masks_random(rng)
        sr.newRegion(image, gamma)
        sr.setOTFMask(m)
        sr.setU(u)
        ll = sr.calcLogLikelihood()
        assert numpy.isclose(ll, numpy.sum(u - (image + gamma) * numpy.log(u + gamma)),
                             rtol=1e-13)
        nc = sr.calcNoiseContribution()
        # py_ref.py:53-58 form: centred mask times |fftshift(fft2 u)| / N, squared
        t = numpy.abs(numpy.fft.fftshift(numpy.fft.fft2(u))) / 16 * m
        assert numpy.isclose(nc, numpy.sum(t ** 2), rtol=1e-12)
    sr.cleanup()


def oracle_masks_random(rng):
    m = numpy.real(numpy.fft.fftshift(numpy.fft.fft2(rng.uniform(0, 1, (16, 16)))))
    m -= m.min()
    return m / m.max()


def test_gradient_central_differences():
    rng = numpy.random.default_rng(5)
    d = rng.uniform(1, 30, (16, 16))
    g = rng.uniform(1, 10, (16, 16))
    u = rng.uniform(1, 30, (16, 16))
    mf = numpy.fft.fftshift(oracle_masks_random(rng))
    f0, grad = oracle.cost_grad(u, d, g, mf, 0.3)
    h = 1e-5
    for idx in [(0, 0), (3, 7), (15, 15), (8, 8), (11, 2)]:
        up, um = u.copy(), u.copy()
        up[idx] += h
        um[idx] -= h
        fd = (oracle.cost_grad(up, d, g, mf, 0.3)[0] - oracle.cost_grad(um, d, g, mf, 0.3)[0]) / (2 * h)
        assert abs(fd - grad[idx]) <= 1e-6 * max(1.0, abs(fd))


def test_fitmin_guard_gradient_follows_the_c():
    """Inside the guard the C returns 1 - c/1e-6 (ncs.c:169-173), not the
    derivative of the guarded log-likelihood."""
    d = numpy.full((16, 16), 5.0)
    g = numpy.full((16, 16), 2.0)
    u = numpy.full((16, 16), 3.0)
    u[4, 5] = -10.0
    mf = numpy.zeros((16, 16))
    f, grad = oracle.cost_grad(u, d, g, mf, 0.0)
    assert grad[4, 5] == 1.0 - 7.0 / 1e-6
    assert numpy.isclose(grad[0, 0], 1.0 - 7.0 / 5.0, rtol=1e-15)
    expect = 255 * (3.0 - 7.0 * numpy.log(5.0)) + (-10.0 - 7.0 * numpy.log(1e-6))
    assert numpy.isclose(f, expect, rtol=1e-13)


def test_parseval_closed_form(golden):
    ones = numpy.ones((16, 16))
    # eps = 1e-9 reaches the FP64 resolution of the cost (|f| ~ 1e4): the
    # line search may end with ROUNDING_ERROR there, at |du| ~ 1e-6.
    p = oracle.default_param(epsilon=1e-9)
    res = oracle.solve_tiles(golden["parseval_d"], golden["parseval_g"], ones,
                             float(golden["parseval_alpha"]), param=p)
    assert set(res["status"]) <= {0, oracle.LBFGSERR_ROUNDING_ERROR}
    assert numpy.max(numpy.abs(res["u"] - golden["parseval_u"])) < 5e-6


def test_alpha_zero_returns_data_unchanged():
    rng = numpy.random.default_rng(7)
    d = rng.uniform(0.5, 50, (4, 16, 16))
    g = rng.uniform(1, 10, (4, 16, 16))
    res = oracle.solve_tiles(d, g, numpy.ones((16, 16)), 0.0)
    assert numpy.all(res["status"] == oracle.LBFGS_ALREADY_MINIMIZED)
    assert numpy.array_equal(res["u"], d)


def test_constant_tile_under_dc_free_mask_is_a_fixed_point():
    mf = numpy.ones((16, 16))
    mf[0, 0] = 0.0  # DC weight zero (FFT order)
    d = numpy.full((1, 16, 16), 17.25)
    g = numpy.full((1, 16, 16), 3.5)
    res = oracle.solve_tiles(d, g, mf, 0.7)
    assert res["status"][0] in (oracle.LBFGS_ALREADY_MINIMIZED, oracle.LBFGS_SUCCESS)
    assert numpy.max(numpy.abs(res["u"][0] - 17.25)) < 1e-9


def test_pure_dc_mask_matches_newton():
    """M nonzero only at DC: NC = m^2 (sum u)^2 / N^2; solved by Newton in FP64."""
    rng = numpy.random.default_rng(11)
    d = rng.uniform(5, 40, (16, 16))
    g = rng.uniform(1, 8, (16, 16))
    mf = numpy.zeros((16, 16))
    mf[0, 0] = 0.5
    alpha = 0.05
    c = (d + g).ravel()
    gv = g.ravel()
    x = d.ravel().copy()
    # NC = m^2 |U_0|^2 / N^2 = 0.25 (sum u)^2 / 256: gradient 2 * 0.25 * sum(u) / 256
    for _ in range(50):
        s = x.sum()
        grad = 1 - c / (x + gv) + 2 * alpha * 0.25 * s / 256.0
        H = numpy.diag(c / (x + gv) ** 2) + 2 * alpha * 0.25 / 256.0
        x = x - numpy.linalg.solve(H, grad)
    res = oracle.solve_tiles(d[None], g[None], mf, alpha, param=oracle.default_param(epsilon=1e-9))
    assert numpy.max(numpy.abs(res["u"][0].ravel() - x)) < 1e-5


@pytest.mark.parametrize("mask", ["otf", "gauss", "random"])
@pytest.mark.parametrize("alpha", [0.02, 0.1])
def test_minimiser_matches_scipy_golden(golden, mask, alpha):
    key = "%s_a%g" % (mask, alpha)
    res = oracle.solve_tiles(golden["solve_d"], golden["solve_g"], golden["mask_" + mask],
                             alpha, param=oracle.default_param(epsilon=1e-9))
    fr = golden["solve_cost_" + key]
    assert numpy.all(numpy.abs(res["cost"] - fr) <= 1e-12 * numpy.abs(fr) + 1e-9)
    assert numpy.max(numpy.abs(res["u"] - golden["solve_u_" + key])) < 1e-4


def test_default_solve_reaches_stop_rule_and_scipy_cost():
    """Reference defaults (m=6, eps=1e-5): stop rule |g| <= eps max(1,|u|)
    holds in an independent numpy gradient; cost within 1e-9 of scipy's."""
    rng = numpy.random.default_rng(13)
    d = rng.uniform(10, 20, (5, 16, 16))
    g = rng.uniform(2, 4, (5, 16, 16))
    mf = numpy.fft.fftshift(oracle_masks_random(rng))
    res = oracle.solve_tiles(d, g, mf, 0.02)
    for i in range(5):
        assert res["status"][i] == 0
        u = res["u"][i]
        v = numpy.maximum(u + g[i], 1e-6)
        gr = 1 - (d[i] + g[i]) / v + 0.02 * 2 * numpy.real(numpy.fft.ifft2(mf ** 2 * numpy.fft.fft2(u)))
        assert numpy.linalg.norm(gr) <= 1e-5 * max(1.0, numpy.linalg.norm(u)) * 1.0001
        ref = scipy.optimize.minimize(lambda x: _np_cost(x.reshape(16, 16), d[i], g[i], mf, 0.02),
                                      d[i].ravel(), method="L-BFGS-B",
                                      options=dict(maxiter=2000, ftol=1e-15, gtol=1e-10))
        assert res["cost"][i] <= ref.fun + 1e-9 * abs(ref.fun)


@pytest.mark.parametrize("r_size", [10, 16])
def test_whole_image_c_tiling_equals_python_tiling(r_size):
    """clib/pyCNCS/test/test_im.py:12-38: the C tiling (ncs.c:71-143) and the
    Python tiling (ncs_c.py:279-322) give the same image at 144 offsets."""
    rng = numpy.random.default_rng(1)
    im_size = 30
    gamma = rng.uniform(2.0, 4.0, (im_size, im_size))
    image = rng.uniform(0.01, 10.0, (im_size, im_size))
    otfmask = oracle_masks_random(numpy.random.default_rng(2))[:r_size, :r_size] \
        if r_size == 16 else _random_mask_r(rng, r_size)
    step = 1 if r_size == 10 else 3
    for ix in range(0, 12, step):
        for iy in range(0, 12, step):
            a = oracle.pyReduceNoise(image[ix:, iy:], gamma[ix:, iy:], otfmask, 0.02)
            b = oracle.cReduceNoise(image[ix:, iy:], gamma[ix:, iy:],
                                    numpy.fft.fftshift(otfmask), 0.02)
            assert numpy.array_equal(a, b), (ix, iy)


def _random_mask_r(rng, r):
    m = numpy.real(numpy.fft.fftshift(numpy.fft.fft2(rng.uniform(0, 1, (r, r)))))
    m -= m.min()
    return m / m.max()


@pytest.mark.parametrize("shape", [(14, 14), (15, 29), (30, 30), (43, 16), (57, 71)])
def test_alpha_zero_whole_image_is_identity(shape):
    rng = numpy.random.default_rng(sum(shape))
    image = rng.uniform(0.5, 100, shape)
    gamma = rng.uniform(1, 10, shape)
    out, info = oracle.reduce_noise(image, gamma, numpy.ones((16, 16)), 0.0)
    assert numpy.array_equal(out, image)


def test_tiles_per_axis_matches_the_loop_of_ncs_c():
    lib = oracle.load()
    for n in range(14, 400):
        cnt, i = 0, -1
        while i < n + 1:
            b = n + 1 - 16 if i + 16 > n + 1 else i
            cnt += 1
            if b == n + 1 - 16:
                break
            i += 14
        assert lib.ncso_tiles_per_axis(n, 16) == cnt
