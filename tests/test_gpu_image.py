"""GPU parity of the whole-image drop-in (tiling + solve + stitch)."""
import ctypes

import numpy
import pytest

import oracle
from ncs_amd import _lib, masks, ncs, sim

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(autouse=True)
def _needs_gpu(gpu_ctx):
    return gpu_ctx


def test_reference_image_test_crops():
    """opencl/pyOpenCLNCS/test/test_im.py:16-44: C vs GPU reduceNoise on
    30 x 30 crops at 144 offsets, r = 16, alpha = 0.02, atol 0.02."""
    _say("reference test_im crops")
    rng = numpy.random.RandomState(1)
    im_size = 30
    gamma = rng.uniform(2.0, 4.0, (im_size, im_size))
    image = rng.uniform(0.01, 10.0, (im_size, im_size))
    otfmask = masks.random_mask(16, numpy.random.default_rng(1))
    worst = 0.0
    for ix in range(12):
        for iy in range(12):
            a = oracle.cReduceNoise(image[ix:, iy:], gamma[ix:, iy:],
                                    numpy.fft.fftshift(otfmask), 0.02)
            b = ncs.reduceNoise([image[ix:, iy:]], gamma[ix:, iy:], otfmask, 0.02)[0]
            worst = max(worst, float(numpy.max(numpy.abs(a - b))))
            assert numpy.allclose(a, b, atol=0.02), (ix, iy)
    _say("worst abs diff over 144 crops", worst)


@pytest.mark.parametrize("shape", [(14, 14), (15, 29), (19, 23), (30, 30), (128, 37), (200, 201)])
def test_alpha_zero_is_bit_exact_identity(shape):
    """alpha = 0 makes every tile's minimiser its data: gather, clamp and
    stitch are then checked bit for bit."""
    _say("identity", shape)
    rng = numpy.random.default_rng(shape[0] * 1000 + shape[1])
    image = rng.uniform(0.5, 100, shape).astype(numpy.float32)
    gamma = rng.uniform(1, 10, shape).astype(numpy.float32)
    out = ncs.reduceNoise([image], gamma, masks.otf_mask(), 0.0)[0]
    assert numpy.array_equal(out, image)
    out_b = ncs.reduceNoise([image], gamma, masks.otf_mask(), 0.0, geometry="blocks")[0]
    assert numpy.array_equal(out_b, image)


def test_parseval_closed_form_image(golden):
    """All-ones mask: per-pixel closed form, whatever the tiling."""
    _say("parseval image")
    rng = numpy.random.default_rng(21)
    d = rng.uniform(5, 60, (61, 77)).astype(numpy.float32)
    g = rng.uniform(1, 20, (61, 77)).astype(numpy.float32)
    a = 0.1
    c = d.astype(numpy.float64) + g
    expect = (-(1 + 2 * a * g) + numpy.sqrt((1 + 2 * a * g) ** 2 - 8 * a * (g - c))) / (4 * a)
    # the stop rule |g| <= eps |u| with Hessian >= 2 alpha bounds the error by
    # eps |u_tile| / (2 alpha) ~ 2.5e-4 (|u_tile| ~ 500) at eps = 1e-7
    p = _lib.default_params(epsilon=1e-7)
    for geom in ("reference", "blocks"):
        out, info = ncs.reduceNoise([d], g, numpy.ones((16, 16)), a, params=p, geometry=geom,
                                    return_info=True)
        err = numpy.max(numpy.abs(out[0] - expect))
        _say(geom, "parseval max abs err", err, "status", numpy.unique(info["status"]))
        assert err < 2.5e-4


def test_synthetic_frame_tile_costs_match_oracle():
    """A 256 x 256 synthetic frame: per-tile final costs (FP64 re-evaluation
    of the GPU tiles) within 1e-5 of the oracle's, stitched images close."""
    _say("synthetic 256x256 frame")
    fr = sim.make_frame(1, (256, 256))
    mask = masks.otf_mask()
    out, info = ncs.reduceNoise([fr["data"]], fr["gamma"], mask, 0.1, return_info=True)
    ref, rinfo = oracle.reduce_noise(fr["data"].astype(numpy.float64),
                                     fr["gamma"].astype(numpy.float64),
                                     masks.to_fft_order(mask), 0.1)
    assert len(info["status"]) == len(rinfo["status"]) == 361
    assert numpy.all(info["status"] == 0)
    rel = numpy.abs(info["cost"].astype(numpy.float64) - rinfo["cost"]) / numpy.abs(rinfo["cost"])
    diff = numpy.abs(out[0] - ref)
    _say("tile cost max rel", rel.max(), "pixel max abs diff", diff.max(),
         "median", numpy.median(diff))
    assert rel.max() <= 1e-5
    assert diff.max() < 0.05 * max(1.0, numpy.max(ref))


def test_uneven_batch_equals_separate_calls():
    """Frames of different sizes, gammas and alphas in one launch give the same
    pixels as one call per frame (each tile's solve is independent)."""
    _say("uneven batch")
    shapes = [(64, 64), (40, 97), (130, 33), (16, 16)]
    rng = numpy.random.default_rng(31)
    frames = [sim.make_frame(100 + i, s) for i, s in enumerate(shapes)]
    alphas = [0.05, 0.1, 0.2, 0.3]
    mask = masks.otf_mask()
    batch = ncs.reduceNoise([f["data"] for f in frames], [f["gamma"] for f in frames],
                            mask, alphas)
    for f, a, b in zip(frames, alphas, batch):
        one = ncs.reduceNoise([f["data"]], f["gamma"], mask, a)[0]
        assert numpy.array_equal(one, b)
    del rng


def test_c_abi_drop_in_matches_python_drop_in():
    """ncsReduceNoise with the clib signature (ncs.h:42) on the GPU."""
    _say("C ABI drop-in")
    fr = sim.make_frame(3, (50, 60))
    mask = masks.otf_mask()
    lib = _lib.load()
    img = fr["data"].astype(numpy.float64)
    gam = fr["gamma"].astype(numpy.float64)
    out = numpy.zeros_like(img)
    mf = numpy.ascontiguousarray(masks.to_fft_order(mask), dtype=numpy.float64)
    lib.ncsReduceNoise(out.ctypes.data, img.ctypes.data, gam.ctypes.data, mf.ctypes.data,
                       ctypes.c_double(0.1), 50, 60, 16)
    py = ncs.reduceNoise([fr["data"]], fr["gamma"], mask, 0.1)[0]
    assert numpy.array_equal(out.astype(numpy.float32), py)
    c = ncs.cReduceNoise(img, gam, mf, 0.1)
    assert numpy.array_equal(c, out)


def test_reference_warnings_and_errors():
    _say("warnings and errors")
    fr = sim.make_frame(4, (32, 32))
    with pytest.raises(ncs.NCSHipException):
        ncs.reduceNoise([fr["data"]], fr["gamma"], numpy.ones((8, 8)), 0.1)
    bad = numpy.random.default_rng(0).uniform(0, 1, (16, 16))
    if not ncs.checkOTFMask(bad):
        with pytest.raises(ncs.NCSHipException):
            ncs.reduceNoise([fr["data"]], fr["gamma"], bad, 0.1)
    with pytest.warns(UserWarning):
        ncs.reduceNoise([fr["data"]], fr["gamma"], masks.otf_mask(), 0.1,
                        params=_lib.default_params(max_iterations=1))


def test_block_geometry_ragged_edges_match_masked_oracle():
    """Non-overlapping 16 x 16 blocks with zero-padded edge blocks whose padded
    pixels are excluded from the likelihood: per-tile final cost against the
    oracle's masked-likelihood solve of the same padded tiles."""
    _say("block geometry, ragged edges")
    fr = sim.make_frame(17, (45, 70))
    mask = masks.otf_mask()
    mf = masks.to_fft_order(mask).astype(numpy.float32)
    out, info = ncs.reduceNoise([fr["data"]], fr["gamma"], mask, 0.15, geometry="blocks",
                                return_info=True)
    H, W = fr["data"].shape
    tiles_d, tiles_g, tiles_l = [], [], []
    for by in range(0, H, 16):
        for bx in range(0, W, 16):
            d = numpy.zeros((16, 16), numpy.float32)
            g = numpy.zeros((16, 16), numpy.float32)
            l = numpy.zeros((16, 16), numpy.float32)
            hh, ww = min(16, H - by), min(16, W - bx)
            d[:hh, :ww] = fr["data"][by:by + hh, bx:bx + ww]
            g[:hh, :ww] = fr["gamma"][by:by + hh, bx:bx + ww]
            l[:hh, :ww] = 1.0
            tiles_d.append(d)
            tiles_g.append(g)
            tiles_l.append(l)
    ref = oracle.solve_tiles_masked(numpy.stack(tiles_d), numpy.stack(tiles_g),
                                    numpy.stack(tiles_l), mf, 0.15)
    assert len(info["cost"]) == len(ref["cost"]) == 3 * 5
    rel = numpy.abs(info["cost"] - ref["cost"]) / numpy.abs(ref["cost"])
    _say("blocks: tile cost max rel", rel.max(), "status", numpy.unique(info["status"]))
    assert numpy.all(info["status"] == 0)
    assert rel.max() <= 1e-5
    # stitched pixels are the valid part of each solved block
    k = 0
    for by in range(0, H, 16):
        for bx in range(0, W, 16):
            hh, ww = min(16, H - by), min(16, W - bx)
            assert numpy.max(numpy.abs(out[0][by:by + hh, bx:bx + ww] - ref["u"][k][:hh, :ww])) < 0.05
            k += 1


def test_pyncs_r14_segmentation_is_the_reference_geometry():
    """pyNCS splits a square frame into R1 x R1 blocks with a one-pixel halo
    (python3-6/pyNCS/denoisetools.py:27-86).  With R1 = 14 its windows are
    the reference geometry's 16 x 16 windows, so solving pyNCS's windows one
    by one and stitching their interiors gives the frame path's image bit
    for bit."""
    from oracle import camera
    _say("pyNCS R1 = 14 segmentation, 112 x 112")
    fr = sim.make_frame(71, (112, 112))
    ctx = _lib.context(0)
    ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    p = _lib.default_params()
    d = camera.segpadimg(fr["data"].astype(numpy.float32), 14).astype(numpy.float32)
    g = camera.segpadimg(fr["gamma"].astype(numpy.float32), 14).astype(numpy.float32)
    u, it, st, co = ctx.solve_tiles(d, g, 0.1, p)
    out = numpy.zeros((112, 112), numpy.float32)
    ctx.reduce_frames([fr["data"].astype(numpy.float32)], [fr["gamma"].astype(numpy.float32)],
                      [out], [0.1], _lib.GEOM_REFERENCE, p)
    seg = camera.stitchpadimg(u)
    _say("max |diff|", float(numpy.max(numpy.abs(seg - out))), "status", numpy.unique(st))
    assert numpy.array_equal(seg.astype(numpy.float32), out)


@pytest.fixture
def _restore_16(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))


@pytest.mark.parametrize("r", [10, 12])
def test_window_side_r_alpha_zero_is_bit_exact_identity(r, _restore_16):
    """Sub-regions of side r (clib ncs.c:41-147 with r_size = r): at alpha = 0
    the image comes back bit for bit, which pins the r-windows' gather,
    clamp and stitch, down to frames of side r - 2."""
    _say("window", r, "identity")
    rng = numpy.random.default_rng(r)
    for shape in [(r - 2, r - 2), (r - 1, 37), (50, 61)]:
        image = rng.uniform(0.5, 100, shape).astype(numpy.float32)
        gamma = rng.uniform(1, 10, shape).astype(numpy.float32)
        mask = masks.to_fft_order(masks.otf_mask(r=r))
        out = ncs.cReduceNoise(image, gamma, mask, 0.0)
        assert numpy.array_equal(out.astype(numpy.float32), image), shape


@pytest.mark.parametrize("r", [10, 12])
def test_window_side_r_parseval_closed_form(r, _restore_16):
    """All-ones r x r mask: NC = sum u^2 in the r-point basis, so every pixel
    has the closed-form minimiser whatever the tiling."""
    _say("window", r, "parseval")
    rng = numpy.random.default_rng(40 + r)
    d = rng.uniform(5, 60, (45, 53)).astype(numpy.float32)
    g = rng.uniform(1, 20, (45, 53)).astype(numpy.float32)
    a = 0.1
    c = d.astype(numpy.float64) + g
    expect = (-(1 + 2 * a * g) + numpy.sqrt((1 + 2 * a * g) ** 2 - 8 * a * (g - c))) / (4 * a)
    h = ncs.NCSHip(size=r, params=_lib.default_params(epsilon=1e-7))
    h.setOTFMask(numpy.ones((r, r)))
    h.setGamma(g)
    out, info = h.reduceNoise([d], a, return_info=True)
    err = float(numpy.max(numpy.abs(out[0] - expect)))
    _say("max abs err", err, "status", numpy.unique(info["status"]))
    assert err < 2.5e-4


@pytest.mark.parametrize("kind", ["otf", "random"])
@pytest.mark.parametrize("r", [10, 12])
def test_window_side_r_tile_costs_match_oracle(r, kind, _restore_16):
    """96 x 96 frames with an r x r mask: per-tile final costs (FP64 oracle
    with the same r) within 1e-5, tile counts equal.  The OTF mask runs on a
    synthetic filament scene at alpha 0.1.  The random point-symmetric mask
    has an odd-odd part, so it exercises the partner coupling of the r-point
    basis (the general-mask build); it runs on the inputs the reference's own
    random-mask test uses (opencl/pyOpenCLNCS/test/test_im.py:16-44: image
    U(0.01, 10), gamma U(2, 4), alpha 0.02).  On bright scenes such masks let
    pixels reach the FITMIN guard, where the reference gradient is not the
    cost's derivative, and the FP32 and FP64 solvers may stop at different
    points."""
    _say("window", r, kind, "oracle")
    if kind == "otf":
        fr = sim.make_frame(90 + r, (96, 96))
        data, gamma, mask, alpha = fr["data"], fr["gamma"], masks.otf_mask(r=r), 0.1
    else:
        rng = numpy.random.default_rng(70 + r)
        data = rng.uniform(0.01, 10.0, (96, 96)).astype(numpy.float32)
        gamma = rng.uniform(2.0, 4.0, (96, 96)).astype(numpy.float32)
        mask, alpha = masks.random_mask(r, numpy.random.default_rng(r)), 0.02
    h = ncs.NCSHip(size=r)
    h.setOTFMask(mask)
    h.setGamma(gamma)
    out, info = h.reduceNoise([data], alpha, return_info=True)
    ref, rinfo = oracle.reduce_noise(data.astype(numpy.float64), gamma.astype(numpy.float64),
                                     masks.to_fft_order(mask), alpha)
    assert len(info["cost"]) == len(rinfo["cost"]) == _lib.tile_count(96, 96, window=r)
    rel = numpy.abs(info["cost"].astype(numpy.float64) - rinfo["cost"]) / numpy.abs(rinfo["cost"])
    _say("tiles", len(rel), "cost max rel", float(rel.max()), "pixel max abs diff",
         float(numpy.max(numpy.abs(out[0] - ref))))
    if kind == "otf":
        assert numpy.all(info["status"] == 0)
    else:
        # -5 (no progress at FP32 precision) is not a failure for the
        # reference either (pyOpenCLNCS/ncs.py:159-162); the cost bar decides
        assert numpy.all((info["status"] == 0) | (info["status"] == -5))
    assert rel.max() <= 1e-5
    # the tile entry point takes tiles of the mask's side
    ctx = _lib.context(0)
    with pytest.raises(ValueError):
        ctx.solve_tiles(numpy.ones((2, 16, 16), numpy.float32),
                        numpy.ones((2, 16, 16), numpy.float32), 0.1, _lib.default_params())
    dt = numpy.stack([data[k:k + r, 2 * k:2 * k + r] for k in range(8)]).astype(numpy.float32)
    gt = numpy.stack([gamma[k:k + r, 2 * k:2 * k + r] for k in range(8)]).astype(numpy.float32)
    ut, itt, stt, ct = ctx.solve_tiles(dt, gt, alpha, _lib.default_params())
    tref = oracle.solve_tiles(dt, gt, masks.to_fft_order(mask).astype(numpy.float32), alpha,
                              f32=True)
    trel = numpy.abs(ct - tref["cost"]) / numpy.abs(tref["cost"])
    _say("r x r tiles: cost max rel", float(trel.max()))
    assert ut.shape == dt.shape and trel.max() <= 1e-5
    with pytest.raises(_lib.NCSHipError):  # the plain start only
        h.params = _lib.default_params(start=_lib.START_LINEAR)
        h.reduceNoise([data], alpha)
