"""GPU checks of the learned linear start fused into the solve kernel.

The start the kernel computes (three basis changes on the matrix cores) must equal the
FP64 numpy predictor of ncs_amd.linstart; solves from it must reach the
plain start's per-tile costs within 1e-5; coefficients that make the start
worse fall back to the plain start per tile."""
import os

import numpy
import pytest

from ncs_amd import _lib, linstart, masks, sim, warmstart

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture
def otf(gpu_ctx):
    gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    return gpu_ctx


def _windows(seed, shape=(96, 96)):
    f = sim.make_frame(seed, shape)
    return (warmstart.reference_windows(f["data"]).astype(numpy.float32),
            warmstart.reference_windows(f["gamma"]).astype(numpy.float32))


def _random_coef(seed, scale=0.3):
    rng = numpy.random.default_rng(seed)
    c = rng.uniform(-scale, scale, linstart.N_COEF).astype(numpy.float32)
    c[512] = 0.25
    return c


def test_kernel_start_matches_numpy_predictor(otf):
    ctx = otf
    d, g = _windows(81)
    coef = _random_coef(1)
    ctx.set_linear_start(coef)
    p = _lib.default_params(start=_lib.START_LINEAR, max_iterations=0, fallback=0)
    u, it, st, co = ctx.solve_tiles(d, g, 0.1, p)
    ref = linstart.predict(d, g, coef)
    err = numpy.abs(u - ref).max(axis=(1, 2))
    scale = numpy.abs(ref).max(axis=(1, 2))
    _say("linear start vs numpy: max err / scale", float((err / scale).max()), "iters", it.max())
    assert numpy.all(it == 0)
    assert numpy.all(err <= 2e-5 * scale)
    # a zero predictor is the plain start itself
    ctx.set_linear_start(numpy.zeros(linstart.N_COEF, numpy.float32))
    u0, _, _, _ = ctx.solve_tiles(d, g, 0.1, p)
    assert numpy.array_equal(u0, d)


def test_kernel_start_follows_alpha_with_slopes(otf):
    """With slopes set, a frame at alpha starts from the predictor at
    coef + ln(alpha / 0.1) * slope (FP64 numpy, 2e-5 of the tile scale); at
    alpha 0.1 itself the start is bit-identical to the one without slopes.
    Slopes need coefficients first."""
    ctx = otf
    d, g = _windows(82)
    coef = _random_coef(3, scale=0.2)
    slope = _random_coef(4, scale=0.1)
    p = _lib.default_params(start=_lib.START_LINEAR, max_iterations=0, fallback=0)
    ctx.set_linear_start(coef)
    u_ref, _, _, _ = ctx.solve_tiles(d, g, 0.1, p)
    ctx.set_linear_start(coef, slope, 0.1)
    u_same, _, _, _ = ctx.solve_tiles(d, g, 0.1, p)
    assert numpy.array_equal(u_same, u_ref)
    for alpha in (0.05, 0.3):
        u, it, _, _ = ctx.solve_tiles(d, g, alpha, p)
        ref = linstart.predict(d, g, coef, alpha=alpha, slope=slope, alpha_ref=0.1)
        err = numpy.abs(u - ref).max(axis=(1, 2))
        scale = numpy.abs(ref).max(axis=(1, 2))
        _say("alpha", alpha, "start vs numpy with slopes: max err / scale",
             float((err / scale).max()))
        assert numpy.all(it == 0)
        assert numpy.all(err <= 2e-5 * scale)
    # alpha <= 0 keeps the coefficients; far alphas are clamped to 8x alpha_ref
    u0a, _, _, _ = ctx.solve_tiles(d, g, 0.0, p)
    assert numpy.all(numpy.isfinite(u0a))
    assert numpy.array_equal(u0a, u_ref)
    u_far, _, _, _ = ctx.solve_tiles(d, g, 50.0, p)
    ref = linstart.predict(d, g, coef, alpha=0.8, slope=slope, alpha_ref=0.1)
    assert numpy.allclose(u_far, ref, rtol=2e-5, atol=2e-5 * float(numpy.abs(ref).max()))
    # set_linear_start alone clears the slopes
    ctx.set_linear_start(coef)
    u_cleared, _, _, _ = ctx.solve_tiles(d, g, 0.3, p)
    assert numpy.allclose(u_cleared, linstart.predict(d, g, coef), rtol=2e-5, atol=1e-3)
    # slopes without coefficients: NCS_E_NOWEIGHTS (-5) on a fresh context
    fresh = _lib.Context(0)
    try:
        rc = fresh.lib.ncs_hip_set_linear_start_alpha(
            fresh.handle, slope.ctypes.data_as(_lib.ctypes.c_void_p), 0.1)
        assert rc == -5
    finally:
        fresh.close()


def test_linear_start_reaches_plain_costs_with_fewer_iterations(otf):
    ctx = otf
    if not os.path.exists(linstart.DEFAULT_COEF):
        pytest.skip("no stored coefficients")
    ctx.set_linear_start(linstart.load())
    d, g = _windows(82, (256, 256))
    _, it_p, st_p, c_p = ctx.solve_tiles(d, g, 0.1, _lib.default_params())
    _, it_l, st_l, c_l = ctx.solve_tiles(d, g, 0.1, _lib.default_params(start=_lib.START_LINEAR))
    rel = numpy.abs(c_l.astype(numpy.float64) - c_p) / numpy.abs(c_p)
    _say("linear vs plain: iterations %.3f -> %.3f, max rel cost %.2e"
         % (it_p.mean(), it_l.mean(), rel.max()))
    assert numpy.all(st_l == 0) and numpy.all(st_p == 0)
    assert rel.max() <= 1e-5
    assert it_l.mean() < it_p.mean() - 0.5


def test_bad_coefficients_fall_back_to_the_plain_start(otf):
    ctx = otf
    d, g = _windows(83)
    coef = numpy.zeros(linstart.N_COEF, numpy.float32)
    coef[512] = 50.0  # shift every pixel by 50 sigma: always worse
    ctx.set_linear_start(coef)
    _, it_p, st_p, c_p = ctx.solve_tiles(d, g, 0.1, _lib.default_params())
    _, it_l, st_l, c_l = ctx.solve_tiles(d, g, 0.1, _lib.default_params(start=_lib.START_LINEAR))
    assert numpy.array_equal(it_l, it_p) and numpy.array_equal(st_l, st_p)
    rel = numpy.abs(c_l.astype(numpy.float64) - c_p) / numpy.abs(c_p)
    assert rel.max() <= 1e-6


def test_linear_start_argument_errors(gpu_ctx):
    with pytest.raises(ValueError):
        gpu_ctx.set_linear_start(numpy.zeros(10, numpy.float32))
    bad = numpy.zeros(linstart.N_COEF, numpy.float32)
    bad[7] = numpy.nan
    with pytest.raises(_lib.NCSHipError):
        gpu_ctx.set_linear_start(bad)
    fresh = _lib.Context(0)
    try:
        d, g = _windows(84, (32, 32))
        with pytest.raises(_lib.NCSHipError):
            fresh.solve_tiles(d, g, 0.1, _lib.default_params(start=_lib.START_LINEAR))
    finally:
        fresh.close()


def test_pipeline_linear_start_matches_eager(otf):
    """The captured pipeline with the linear start gives the eager batch
    entry's image and costs bit for bit."""
    import torch
    ctx = otf
    ctx.set_linear_start(_random_coef(2, scale=0.1))
    shape = (80, 120)
    f = sim.make_frame(85, shape)
    p = _lib.default_params(start=_lib.START_LINEAR)
    pl = _lib.Pipeline(ctx, [shape], [0.1], params=p, keep_tile_results=True)
    pl.set_gamma(0, f["gamma"])
    src = torch.zeros(shape, dtype=torch.float32, pin_memory=True).numpy()
    out = torch.zeros(shape, dtype=torch.float32, pin_memory=True).numpy()
    src[...] = f["data"]
    for _ in range(2):
        pl.launch([src], [out])
        pl.wait()
        it_p, st_p, co_p = pl.tile_results()
        ref = numpy.zeros(shape, numpy.float32)
        it, st, co = ctx.reduce_frames([f["data"]], [f["gamma"]], [ref], [0.1],
                                       _lib.GEOM_REFERENCE, p)
        assert numpy.array_equal(out, ref)
        assert numpy.array_equal(co_p, co) and numpy.array_equal(it_p, it)
    pl.close()
