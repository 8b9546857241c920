"""GPU checks of the solver's precision machinery.

By default every basis change runs on FP32 MFMA.  With
precision=PRECISION_F16_SPLIT they run on the FP16 matrix cores with every
operand split into an FP16 hi/lo pair.  Data-like operands carry a per-tile power-of-two
scale, a non-finite result falls back to FP32, and past a 1e-5 relative
gradient the solver switches to FP32.  Data over nine decades of scale
must still reach the FP64 oracle's minimiser, to the bars of
test_gpu_solver.py: final cost within 1e-5 relative, and the FP64 gradient
within the stop rule.  Overflow is forced in the two transforms that can
meet it: the learned start's gamma transform and the likelihood-gradient
transform.
"""
import numpy
import pytest

import oracle
from ncs_amd import _lib, masks, sim

pytestmark = pytest.mark.gpu

F16 = _lib.PRECISION_F16_SPLIT
FP32 = _lib.PRECISION_FP32


def _say(*a):
    print(*a, flush=True)


def _tiles(seed, n):
    fr = sim.make_frame(seed, (128, 128))
    rng = numpy.random.default_rng(seed)
    ys = rng.integers(0, 112, n)
    xs = rng.integers(0, 112, n)
    d = numpy.stack([fr["data"][y:y + 16, x:x + 16] for y, x in zip(ys, xs)])
    g = numpy.stack([fr["gamma"][y:y + 16, x:x + 16] for y, x in zip(ys, xs)])
    return d.astype(numpy.float64), g.astype(numpy.float64)


def _check(ctx, d, g, params, label, alpha=0.1, mask=None):
    m = masks.otf_mask() if mask is None else mask
    mf = masks.to_fft_order(m).astype(numpy.float32)
    ctx.set_otf_mask(mf)
    d32, g32 = d.astype(numpy.float32), g.astype(numpy.float32)
    u, it, st, cost = ctx.solve_tiles(d32, g32, alpha, params)
    ref = oracle.solve_tiles(d32, g32, mf, alpha, f32=True)
    f64, gn, un = oracle.eval_cost_f32(u, d32, g32, mf, alpha)
    rel = (f64 - ref["cost"]) / numpy.abs(ref["cost"])
    _say(label, "max rel cost", float(rel.max()), "iters", float(it.mean()),
         "status", dict(zip(*[x.tolist() for x in numpy.unique(st, return_counts=True)])),
         "max |g|/max(1,|u|)", float(numpy.max(gn / numpy.maximum(1.0, un))))
    assert numpy.all(st == 0), label
    assert rel.max() <= 1e-5, label
    assert numpy.all(gn <= 2e-5 * numpy.maximum(1.0, un)), label
    return it


@pytest.mark.parametrize("precision", [FP32, F16])
@pytest.mark.parametrize("scale", [1e-3, 1.0, 1e4, 1e6])
def test_scaled_data_reach_the_oracle_minimiser(gpu_ctx, scale, precision):
    """Photon data times 1e-3 .. 1e6: the per-tile scale keeps the FP16
    operands in range and precise at every magnitude; FP32 needs no scale."""
    d, g = _tiles(11, 64)
    _check(gpu_ctx, d * scale, g * scale, _lib.default_params(precision=precision),
           "precision %d scale %g" % (precision, scale))


def _rotated_mask(angle=0.5, sa=3.0, sb=6.0):
    """1 - an anisotropic Gaussian low-pass rotated by `angle`, made exactly
    point-symmetric (m(k) = m(-k) on the 16-point grid, Nyquist included);
    centred 16 x 16 layout, values in [0, 1]."""
    k = numpy.fft.fftfreq(16, 1.0 / 16)
    ky, kx = numpy.meshgrid(k, k, indexing="ij")
    a = numpy.cos(angle) * kx + numpy.sin(angle) * ky
    b = -numpy.sin(angle) * kx + numpy.cos(angle) * ky
    m = 1.0 - numpy.exp(-(a / sa) ** 2 - (b / sb) ** 2)  # FFT order
    neg = (-numpy.arange(16)) % 16
    m = 0.5 * (m + m[neg][:, neg])
    return numpy.fft.fftshift(m)


@pytest.mark.parametrize("alpha", [2.0, 20.0])
def test_strong_regularisation_still_meets_the_stop_rule(gpu_ctx, alpha):
    """The FP32 build accepts convergence on its incremental NC gradient when
    the gradient's rounding-drift bound, which grows with alpha, cannot
    cross the stop rule, and confirms with an exact gradient otherwise.  At
    large alpha the bound exceeds the margin: the FP64 gradient must still
    meet the stop rule and the cost the oracle's."""
    d, g = _tiles(17, 64)
    # the OTF mask, and a rotated anisotropic one: point-symmetric but not
    # even along each axis, so its odd-odd part couples coefficient pairs
    # (the general build, whose |K| bound includes the coupling)
    for name, mask in (("otf", None), ("rotated", _rotated_mask())):
        for precision in (FP32, F16):
            it = _check(gpu_ctx, d, g, _lib.default_params(precision=precision),
                        "alpha %g %s mask precision %d" % (alpha, name, precision),
                        alpha=alpha, mask=mask)
            if name == "rotated":
                # H0 inverts the coupled 2 x 2 blocks: about 3.5 iterations
                # (12-17 with the diagonal alone)
                assert it.mean() <= 6.0, (name, alpha, precision, it.mean())


@pytest.mark.parametrize("history", [1, 3])
@pytest.mark.parametrize("precision", [FP32, F16])
def test_corrected_h0_with_coupled_mask_coefficients(gpu_ctx, history, precision):
    """The curvature-corrected initial inverse Hessian under the rotated
    anisotropic mask at alpha 0.1 (the ratio guard keeps the correction on):
    its P inverts the 2 x 2 blocks the odd-odd part couples.  Costs and the
    FP64 stop rule against the oracle as for every mask, and fewer iterations
    than the spectral H0 on average."""
    d, g = _tiles(19, 96)
    mask = _rotated_mask()
    its = {}
    for pc in (1, 2):
        its[pc] = _check(gpu_ctx, d, g,
                         _lib.default_params(precision=precision, history=history, precond=pc),
                         "rotated mask, history %d, precision %d, precond %d"
                         % (history, precision, pc), alpha=0.1, mask=mask).mean()
    _say("rotated mask: iterations precond 1 %.2f, precond 2 %.2f" % (its[1], its[2]))
    assert its[2] < its[1] - 0.3, its


def test_overflowing_gamma_transform_falls_back_to_fp32(gpu_ctx):
    """gamma 1e4 times the data puts the learned start's gamma transform far
    past FP16's range under the tile's data scale: the FP32 fallback must
    produce the start the FP64 predictor gives."""
    from ncs_amd import linstart
    d, g = _tiles(14, 32)
    d32 = d.astype(numpy.float32)
    g32 = (1e4 * d.mean() * (g / g.mean())).astype(numpy.float32)
    gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    coef = numpy.random.default_rng(3).uniform(-0.3, 0.3, linstart.N_COEF).astype(numpy.float32)
    gpu_ctx.set_linear_start(coef)
    p = _lib.default_params(start=_lib.START_LINEAR, max_iterations=0, fallback=0,
                            precision=F16)
    u, it, st, cost = gpu_ctx.solve_tiles(d32, g32, 0.1, p)
    ref = linstart.predict(d32, g32, coef)
    err = numpy.abs(u - ref).max(axis=(1, 2)) / numpy.abs(ref).max(axis=(1, 2))
    _say("gamma 1e4 x data: start vs predictor max rel", float(err.max()))
    assert numpy.all(numpy.isfinite(u)) and err.max() <= 2e-5


def test_overflowing_gradient_falls_back_to_fp32(gpu_ctx):
    """A start 1e5 below the data with gamma near 0 makes the likelihood
    gradient about -1e5 per pixel, past FP16's range in the (unscaled)
    gradient transform: the FP32 fallback must give a finite first step
    that lowers the cost of every tile."""
    d, g = _tiles(12, 32)
    d32 = d.astype(numpy.float32)
    g32 = numpy.full_like(d32, 1e-6)
    u0 = (1e-5 * d).astype(numpy.float32)
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    import torch
    dev = torch.device("cuda:0")
    td, tg, tu0 = (torch.from_numpy(x).to(dev) for x in (d32, g32, u0))
    tu = torch.empty_like(td)
    ti = torch.empty(len(d), dtype=torch.int32, device=dev)
    ts = torch.empty_like(ti)
    p = _lib.default_params(fallback=0, max_iterations=1, precision=F16)
    gpu_ctx.solve_tiles_device(td.data_ptr(), tg.data_ptr(), tu0.data_ptr(), tu.data_ptr(),
                               ti.data_ptr(), ts.data_ptr(), None, len(d), 0.1, p,
                               torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    u1 = tu.cpu().numpy()
    f0, _, _ = oracle.eval_cost_f32(u0, d32, g32, mf, 0.1)
    f1, _, _ = oracle.eval_cost_f32(u1, d32, g32, mf, 0.1)
    _say("overflowing start: status", numpy.unique(ts.cpu().numpy()),
         "cost ratio max", float(numpy.max(f1 / f0)))
    assert numpy.all(numpy.isfinite(u1))
    assert numpy.all(f1 < f0)


def test_tight_tolerance_finishes_in_fp32(gpu_ctx):
    """eps = 1e-7 is past the FP16 split's accuracy: the solver must switch
    to FP32 and meet the tighter stop rule."""
    d, g = _tiles(13, 32)
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    d32, g32 = d.astype(numpy.float32), g.astype(numpy.float32)
    u, it, st, cost = gpu_ctx.solve_tiles(d32, g32, 0.1,
                                          _lib.default_params(epsilon=1e-7, precision=F16))
    f64, gn, un = oracle.eval_cost_f32(u, d32, g32, mf, 0.1)
    _say("eps 1e-7: status", numpy.unique(st), "max |g|/max(1,|u|)",
         float(numpy.max(gn / numpy.maximum(1.0, un))))
    assert numpy.all(st == 0)
    assert numpy.all(gn <= 5e-7 * numpy.maximum(1.0, un))


def test_stalled_fp16_search_retries_in_fp32(gpu_ctx):
    """A bright scene under a random 10 x 10 mask holds a tile whose line
    search, in FP16 basis changes, stalls far from the minimum (the
    likelihood and NC parts of phi'(t) nearly cancel).  The tile must retry
    in FP32 and reach the FP64 oracle's cost like every other tile."""
    from ncs_amd import ncs
    r = 10
    fr = sim.make_frame(100, (96, 96))
    mask = masks.random_mask(r, numpy.random.default_rng(r))
    try:
        h = ncs.NCSHip(size=r, params=_lib.default_params(precision=F16))
        h.setOTFMask(mask)
        h.setGamma(fr["gamma"])
        out, info = h.reduceNoise([fr["data"]], 0.1, return_info=True)
    finally:
        gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    ref, rinfo = oracle.reduce_noise(fr["data"].astype(numpy.float64),
                                     fr["gamma"].astype(numpy.float64),
                                     masks.to_fft_order(mask), 0.1)
    rel = numpy.abs(info["cost"].astype(numpy.float64) - rinfo["cost"]) / numpy.abs(rinfo["cost"])
    _say("random 10x10 mask, bright scene: status",
         dict(zip(*[x.tolist() for x in numpy.unique(info["status"], return_counts=True)])),
         "max rel cost", float(rel.max()))
    assert numpy.all(info["status"] == 0)
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("precision", [FP32, F16])
def test_failed_corrected_search_restarts_from_the_data(gpu_ctx, precision):
    """A line search that fails under the curvature-corrected initial inverse
    Hessian restarts the tile from the plain start with the spectral one.  On
    this bright scene under a random 12 x 12 mask (history 1, alpha 0.1) the
    spectral H0 alone ends 2-3 tiles in failed searches and a retry from the
    failing point 1; with the restart every tile converges.  Costs: within
    1e-5 of the FP64 oracle where it converged, no higher where its
    liblbfgs restatement stopped early (random masks make some of these
    tiles stiff for it too)."""
    from ncs_amd import ncs
    r = 12
    fr = sim.make_frame(101, (96, 96))
    mask = masks.random_mask(r, numpy.random.default_rng(r))
    try:
        h = ncs.NCSHip(size=r, params=_lib.default_params(precision=precision, history=1))
        h.setOTFMask(mask)
        h.setGamma(fr["gamma"])
        out, info = h.reduceNoise([fr["data"]], 0.1, return_info=True)
    finally:
        gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    ref, rinfo = oracle.reduce_noise(fr["data"].astype(numpy.float64),
                                     fr["gamma"].astype(numpy.float64),
                                     masks.to_fft_order(mask), 0.1)
    rel = (info["cost"].astype(numpy.float64) - rinfo["cost"]) / numpy.abs(rinfo["cost"])
    conv = numpy.isin(rinfo["status"], (0, 2))
    _say("random 12x12 mask, history 1, precision", precision, "status",
         dict(zip(*[x.tolist() for x in numpy.unique(info["status"], return_counts=True)])),
         "oracle converged", int(conv.sum()), "of", conv.size,
         "max rel cost", float(numpy.abs(rel[conv]).max()))
    assert numpy.all(info["status"] == 0)
    assert numpy.abs(rel[conv]).max() <= 1e-5
    assert numpy.all(rel[~conv] <= 1e-5)
