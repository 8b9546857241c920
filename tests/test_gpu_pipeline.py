"""GPU tests of the captured streaming pipeline and of the autotuner.

A replay of the captured graph must give bit for bit what the eager batch
entry gives on the same inputs: both run the same per-tile solver, and a
tile's result does not depend on which wave or schedule solved it.  Every
replay is checked, with eager work and synchronisations between replays.
"""
import numpy
import pytest

from ncs_amd import _lib, masks, sim

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


class _Pinned(object):
    """Pinned host arrays (kept alive by this object)."""

    def __init__(self):
        self.keep = []

    def __call__(self, shape, dtype):
        import torch
        tdt = {numpy.dtype(numpy.float32): torch.float32,
               numpy.dtype(numpy.uint16): torch.int16}[numpy.dtype(dtype)]
        t = torch.zeros(shape, dtype=tdt, pin_memory=True)
        self.keep.append(t)
        return t.numpy().view(dtype)


def _frames(seeds_shapes):
    out = []
    for seed, shape in seeds_shapes:
        f = sim.make_frame(seed, shape)
        out.append(f)
    return out


def _eager(ctx, datas, gammas, alphas, params, geometry=_lib.GEOM_REFERENCE):
    outs = [numpy.zeros(d.shape, numpy.float32) for d in datas]
    it, st, co = ctx.reduce_frames(datas, gammas, outs, alphas, geometry, params)
    return outs, it, st, co


@pytest.fixture
def otf(gpu_ctx):
    gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    return gpu_ctx


def test_pipeline_replays_match_eager_and_follow_new_buffers(otf):
    ctx = otf
    pin = _Pinned()
    fr = _frames([(31, (150, 203)), (32, (64, 96)), (33, (150, 203))])
    alphas = [0.1, 0.25, 0.05]
    shapes = [f["data"].shape for f in fr]
    p = _lib.default_params()
    pl = _lib.Pipeline(ctx, shapes, alphas, params=p, keep_tile_results=True)
    for i, f in enumerate(fr):
        pl.set_gamma(i, f["gamma"])
    ins_a = [pin(s, numpy.float32) for s in shapes]
    ins_b = [pin(s, numpy.float32) for s in shapes]
    outs_a = [pin(s, numpy.float32) for s in shapes]
    outs_b = [pin(s, numpy.float32) for s in shapes]
    rng = numpy.random.default_rng(5)
    for i, f in enumerate(fr):
        ins_a[i][...] = f["data"]
        ins_b[i][...] = numpy.maximum(f["data"] * rng.uniform(0.8, 1.2, f["data"].shape),
                                      1e-6).astype(numpy.float32)
    gam = [f["gamma"] for f in fr]
    ref_a = _eager(ctx, [a.copy() for a in ins_a], gam, alphas, p)
    ref_b = _eager(ctx, [a.copy() for a in ins_b], gam, alphas, p)
    plan = [("a", ins_a, outs_a), ("b", ins_b, outs_b), ("a", ins_a, outs_b),
            ("a", None, None), ("b", ins_b, outs_a)]
    for k, (which, ins, outs) in enumerate(plan):
        _say("replay", k, which)
        pl.launch(ins, outs)
        pl.wait()
        got = outs if outs is not None else outs_b
        ref = ref_a if which == "a" else ref_b
        for i in range(len(fr)):
            assert numpy.array_equal(got[i], ref[0][i]), (k, i)
        it, st, co = pl.tile_results()
        assert numpy.array_equal(it, ref[1]) and numpy.array_equal(st, ref[2])
        assert numpy.array_equal(co, ref[3])
        # eager work between replays
        _eager(ctx, [fr[1]["data"]], [fr[1]["gamma"]], [0.3], p)
    pl.close()


def test_pipeline_raw_adu_input_converts_on_the_device(otf):
    """uint16 ADU frames with camera maps: alpha = 0 returns the photon image
    itself, which must equal (ADU - offset) / gain in float32 with values <= 0
    set to 1e-6 (denoisetools.py:218-224) bit for bit; gamma = var / gain^2."""
    ctx = otf
    pin = _Pinned()
    shape = (77, 130)
    f = sim.make_frame(41, shape)
    adu = numpy.clip(numpy.rint(f["adu"]), 0, 65535).astype(numpy.uint16)
    off = f["offset"].astype(numpy.float32)
    gain = f["gain"].astype(numpy.float32)
    var = f["var"].astype(numpy.float32)
    adu[0, :5] = 0  # values <= 0 after the offset
    photons = (adu.astype(numpy.float32) - off) / gain
    photons[photons <= 0] = numpy.float32(1e-6)
    gamma = var / (gain * gain)
    p = _lib.default_params()
    for alpha in (0.0, 0.1):
        pl = _lib.Pipeline(ctx, [shape], [alpha], params=p, input_format=_lib.INPUT_ADU_U16,
                           keep_tile_results=True)
        pl.set_camera(0, off, gain, var)
        src = pin(shape, numpy.uint16)
        src[...] = adu
        out = pin(shape, numpy.float32)
        pl.launch([src], [out])
        pl.wait()
        ref = _eager(ctx, [photons], [gamma], [alpha], p)
        _say("adu alpha", alpha, "max |diff|", float(numpy.max(numpy.abs(out - ref[0][0]))))
        if alpha == 0.0:
            assert numpy.array_equal(out, photons)
        assert numpy.array_equal(out, ref[0][0])
        assert numpy.array_equal(pl.tile_results()[2], ref[3])
        pl.close()


def test_pipeline_gain_levels_pick_maps_by_adu(otf):
    """Three gain amplifiers (matlab/genmaps.m:29-62): each pixel's gain and
    variance maps are picked by its ADU against the division points, ties
    included, and non-integer points round as ADU >= point does.  alpha = 0
    returns the photon image, bit for bit the oracle's genmaps + conversion;
    alpha = 0.1 gives the eager batch's image and tile results on those
    photons and gammas.  New points recapture the graph; one level again is
    the single-map conversion."""
    from oracle import camera
    ctx = otf
    pin = _Pinned()
    shape = (77, 130)
    f = sim.make_frame(42, shape)
    adu = numpy.clip(numpy.rint(f["adu"]), 0, 65535).astype(numpy.uint16)
    off = f["offset"].astype(numpy.float32)
    gains = numpy.stack([f["gain"], 1.3 * f["gain"], 2.1 * f["gain"]]).astype(numpy.float32)
    var = numpy.stack([f["var"], 4.0 * f["var"], 9.0 * f["var"]]).astype(numpy.float32)
    lo, hi = (int(v) for v in numpy.percentile(adu, [40, 80]))
    adu[0, :4] = [lo - 1, lo, hi - 1, hi]  # ties at the division points
    src = pin(shape, numpy.uint16)
    src[...] = adu
    out = pin(shape, numpy.float32)
    p = _lib.default_params()

    def expect(points):
        gs, vs = camera.genmaps(adu[None], gains, var, points)
        return (camera.adu_to_photons(adu, off, gs[0]),
                camera.gamma_from_camera(vs[0], gs[0]))

    for alpha in (0.0, 0.1):
        pl = _lib.Pipeline(ctx, [shape], [alpha], params=p, input_format=_lib.INPUT_ADU_U16,
                           keep_tile_results=True)
        for points in ([lo, hi], [lo - 0.5, hi + 0.25]):
            pl.set_camera_levels(0, points, off, gains, var)
            ph, gm = expect(points)
            pl.launch([src], [out])
            pl.wait()
            ref = _eager(ctx, [ph], [gm], [alpha], p)
            _say("levels", points, "alpha", alpha, "max |diff|",
                 float(numpy.max(numpy.abs(out - ref[0][0]))))
            if alpha == 0.0:
                assert numpy.array_equal(out, ph)
            assert numpy.array_equal(out, ref[0][0])
            assert numpy.array_equal(pl.tile_results()[2], ref[3])
        with pytest.raises(_lib.NCSHipError):
            pl.set_gamma(0, gm)  # gamma comes from the levels
        pl.set_camera(0, off, gains[1], var[1])
        pl.launch([src], [out])
        pl.wait()
        ph1 = camera.adu_to_photons(adu, off, gains[1])
        ref = _eager(ctx, [ph1], [camera.gamma_from_camera(var[1], gains[1])], [alpha], p)
        assert numpy.array_equal(out, ref[0][0])
        pl.close()
    pl = _lib.Pipeline(ctx, [shape], [0.1], params=p, input_format=_lib.INPUT_ADU_U16)
    with pytest.raises(_lib.NCSHipError):
        pl.set_camera_levels(0, [hi, lo], off, gains, var)  # descending
    with pytest.raises(_lib.NCSHipError):
        five = numpy.concatenate([gains, gains[:2]])
        pl.set_camera_levels(0, [1, 2, 3, 4], off, five, numpy.concatenate([var, var[:2]]))
    with pytest.raises(ValueError):
        pl.set_camera_levels(0, [lo], off, gains, var)  # 3 levels take 2 points
    pl.close()


def test_pipeline_gain_levels_per_frame(otf):
    """Two frames of different sizes in one pipeline, the first with two
    gain levels and the second with one map: each frame's photons and gamma
    come from its own maps (the level maps sit at each frame's offset in the
    packed buffers)."""
    from oracle import camera
    ctx = otf
    pin = _Pinned()
    shapes = [(40, 50), (33, 64)]
    fr = [sim.make_frame(45 + i, s) for i, s in enumerate(shapes)]
    p = _lib.default_params()
    pl = _lib.Pipeline(ctx, shapes, [0.1, 0.2], params=p, input_format=_lib.INPUT_ADU_U16,
                       keep_tile_results=True)
    srcs, outs, phs, gms = [], [], [], []
    for i, f in enumerate(fr):
        adu = numpy.clip(numpy.rint(f["adu"]), 0, 65535).astype(numpy.uint16)
        off = f["offset"].astype(numpy.float32)
        if i == 0:
            gains = numpy.stack([f["gain"], 1.7 * f["gain"]]).astype(numpy.float32)
            var = numpy.stack([f["var"], 3.0 * f["var"]]).astype(numpy.float32)
            pt = [float(numpy.median(adu))]
            pl.set_camera_levels(0, pt, off, gains, var)
            gs, vs = camera.genmaps(adu[None], gains, var, pt)
            g0, v0 = gs[0], vs[0]
        else:
            g0, v0 = f["gain"].astype(numpy.float32), f["var"].astype(numpy.float32)
            pl.set_camera(1, off, g0, v0)
        phs.append(camera.adu_to_photons(adu, off, g0))
        gms.append(camera.gamma_from_camera(v0, g0))
        srcs.append(pin(shapes[i], numpy.uint16))
        srcs[-1][...] = adu
        outs.append(pin(shapes[i], numpy.float32))
    ref = _eager(ctx, phs, gms, [0.1, 0.2], p)
    for k in range(2):
        pl.launch(srcs, outs)
        pl.wait()
        for i in range(2):
            assert numpy.array_equal(outs[i], ref[0][i]), (k, i)
        assert numpy.array_equal(pl.tile_results()[2], ref[3])
    pl.close()


def test_pipeline_blocks_geometry_uneven_frames(otf):
    ctx = otf
    pin = _Pinned()
    shapes = [(64, 64), (100, 37), (17, 250)]
    alphas = [0.05, 0.3, 0.17]
    fr = _frames([(51 + i, s) for i, s in enumerate(shapes)])
    p = _lib.default_params()
    pl = _lib.Pipeline(ctx, shapes, alphas, params=p, geometry=_lib.GEOM_BLOCKS)
    for i, f in enumerate(fr):
        pl.set_gamma(i, f["gamma"])
    ins = [pin(s, numpy.float32) for s in shapes]
    outs = [pin(s, numpy.float32) for s in shapes]
    for i, f in enumerate(fr):
        ins[i][...] = f["data"]
    ref = _eager(ctx, [f["data"] for f in fr], [f["gamma"] for f in fr], alphas, p,
                 geometry=_lib.GEOM_BLOCKS)
    for k in range(3):
        pl.launch(ins, outs)
        pl.wait()
        for i in range(len(shapes)):
            assert numpy.array_equal(outs[i], ref[0][i]), (k, i)
    pl.close()


def test_pipeline_rejects_missing_gamma_and_bad_buffers(otf):
    pl = _lib.Pipeline(otf, [(32, 32)], [0.1], keep_tile_results=True)
    a = numpy.zeros((32, 32), numpy.float32)
    with pytest.raises(_lib.NCSHipError):
        pl.launch([a], [a.copy()])  # gamma not set
    pl.set_gamma(0, numpy.ones((32, 32), numpy.float32))
    with pytest.raises(ValueError):
        pl.launch([numpy.zeros((31, 32), numpy.float32)], [a])
    with pytest.raises(_lib.NCSHipError):
        pl.launch(None, None)  # the first launch needs buffers
    with pytest.raises(_lib.NCSHipError):
        pl.tile_results()  # nothing launched
    with pytest.raises(_lib.NCSHipError):
        pl.wait(0)
    assert pl.launched() == 0
    pl.close()


def test_replays_in_flight_on_two_streams_match_eager(otf):
    """Seven replays enqueued without host waits, alternating between the
    pipeline's stream and a second one, each with its own input and output:
    the device-side ordering must keep every replay's image and tile results
    those of the eager batch on its input.  Seven is past the four replay
    slots, so the launch that reuses a slot waits for its previous replay."""
    import torch
    ctx = otf
    pin = _Pinned()
    shape = (96, 130)
    f = sim.make_frame(61, shape)
    p = _lib.default_params()
    n = 7
    rng = numpy.random.default_rng(8)
    ins = [pin(shape, numpy.float32) for _ in range(n)]
    outs = [pin(shape, numpy.float32) for _ in range(n)]
    refs = []
    for k in range(n):
        ins[k][...] = numpy.maximum(f["data"] * rng.uniform(0.5, 1.5, shape), 1e-6)
        refs.append(_eager(ctx, [ins[k].copy()], [f["gamma"]], [0.1], p))
    pl = _lib.Pipeline(ctx, [shape], [0.1], params=p, keep_tile_results=True)
    pl.set_gamma(0, f["gamma"])
    other = torch.cuda.Stream()
    seqs = [pl.launch([ins[k]], [outs[k]], None if k % 2 == 0 else other.cuda_stream)
            for k in range(n)]
    assert seqs == list(range(n)) and pl.launched() == n
    _say("launched", n, "replays")
    for k in range(n - _lib.PIPELINE_REPLAYS, n):
        it, st, co = pl.tile_results(k)
        assert numpy.array_equal(co, refs[k][3]) and numpy.array_equal(it, refs[k][1]), k
        assert numpy.array_equal(st, refs[k][2]), k
    with pytest.raises(_lib.NCSHipError):
        pl.tile_results(n - _lib.PIPELINE_REPLAYS - 1)  # its slot was reused
    pl.wait()
    for k in range(n):
        assert numpy.array_equal(outs[k], refs[k][0][0]), k
    # a replay with new outputs only keeps the last inputs
    out2 = pin(shape, numpy.float32)
    seq = pl.launch(None, [out2])
    pl.wait(seq)
    assert numpy.array_equal(out2, refs[n - 1][0][0])
    pl.close()


def _device_frames(datas, gammas, alphas, geometry=_lib.GEOM_REFERENCE):
    import torch
    dev = torch.device("cuda:0")
    keep, frames = [], []
    for d, g, a in zip(datas, gammas, alphas):
        td = torch.from_numpy(numpy.ascontiguousarray(d)).to(dev)
        tg = torch.from_numpy(numpy.ascontiguousarray(g)).to(dev)
        to = torch.zeros_like(td)
        keep += [td, tg, to]
        frames.append(dict(data=td.data_ptr(), gamma=tg.data_ptr(), out=to.data_ptr(),
                           height=d.shape[0], width=d.shape[1], alpha=a, geometry=geometry))
    return frames, keep


def test_every_schedule_gives_identical_tiles(otf):
    """Workgroup size and grid change who solves a tile, never its result."""
    import torch
    ctx = otf
    f = sim.make_frame(61, (160, 160))
    frames, keep = _device_frames([f["data"]], [f["gamma"]], [0.1])
    n = _lib.tile_count(160, 160)
    res = {}
    for wpg in (1, 2, 4):
        for wpc in (0, 3):
            cost = torch.zeros(n, dtype=torch.float32, device="cuda:0")
            p = _lib.default_params(waves_per_group=wpg, waves_per_cu=wpc)
            ctx.reduce_frames_device(frames, p, cost_ptr=cost.data_ptr())
            torch.cuda.synchronize()
            res[(wpg, wpc)] = (cost.cpu().numpy(), keep[2].cpu().numpy())
    base = res[(1, 0)]
    for k, v in res.items():
        assert numpy.array_equal(v[0], base[0]) and numpy.array_equal(v[1], base[1]), k


def test_autotune_picks_a_schedule_that_is_used(otf):
    import torch
    ctx = otf
    f = sim.make_frame(62, (512, 512))
    frames, keep = _device_frames([f["data"]], [f["gamma"]], [0.1])
    r = ctx.autotune(frames, _lib.default_params(), tune_history=True)
    _say("autotune", r)
    assert r["waves_per_group"] in (1, 2, 4)
    assert 1 <= r["waves_per_cu"] <= r["lds_waves_per_cu"]
    assert r["history"] in (1, 2, 3, 5, 8) and r["candidates"] >= 10
    assert 0 < r["ms"] <= r["ms_default"]
    # nothing written to the sample's output
    assert float(keep[2].abs().max()) == 0.0
    n = _lib.tile_count(512, 512)
    cost = {}
    for name, p in (("tuned", _lib.default_params(history=0)),
                    ("m5", _lib.default_params(history=5))):
        c = torch.zeros(n, dtype=torch.float32, device="cuda:0")
        st = torch.zeros(n, dtype=torch.int32, device="cuda:0")
        ctx.reduce_frames_device(frames, p, cost_ptr=c.data_ptr(), status_ptr=st.data_ptr())
        torch.cuda.synchronize()
        assert int((st != 0).sum()) == 0, name
        cost[name] = c.cpu().numpy().astype(numpy.float64)
    rel = numpy.abs(cost["tuned"] - cost["m5"]) / numpy.abs(cost["m5"])
    _say("tuned vs history 5: max rel cost diff", rel.max())
    assert rel.max() <= 1e-5


def test_pipeline_learned_start_matches_eager_learned(otf):
    """The net's start inside the captured graph gives what the eager
    device entry gives with NCS_START_LEARNED, bit for bit, and the costs
    match the plain start's within 1e-5."""
    import torch
    from ncs_amd import warmstart
    ctx = otf
    warmstart.set_context_weights(warmstart.load())
    pin = _Pinned()
    shape = (96, 140)
    f = sim.make_frame(71, shape)
    p_l = _lib.default_params(start=_lib.START_LEARNED)
    pl = _lib.Pipeline(ctx, [shape], [0.1], params=p_l, keep_tile_results=True)
    pl.set_gamma(0, f["gamma"])
    src = pin(shape, numpy.float32)
    src[...] = f["data"]
    out = pin(shape, numpy.float32)
    pl.launch([src], [out])
    pl.wait()
    it_p, st_p, co_p = pl.tile_results()
    pl.close()
    frames, keep = _device_frames([f["data"]], [f["gamma"]], [0.1])
    n = _lib.tile_count(*shape)
    cost = torch.zeros(n, dtype=torch.float32, device="cuda:0")
    ctx.reduce_frames_device(frames, p_l, cost_ptr=cost.data_ptr())
    torch.cuda.synchronize()
    assert numpy.array_equal(out, keep[2].cpu().numpy())
    assert numpy.array_equal(co_p, cost.cpu().numpy())
    plain = _eager(ctx, [f["data"]], [f["gamma"]], [0.1], _lib.default_params())
    rel = numpy.abs(co_p.astype(numpy.float64) - plain[3]) / numpy.abs(plain[3])
    _say("pipeline learned vs plain: max rel cost", rel.max(), "iters", it_p.mean(), plain[1].mean())
    assert rel.max() <= 1e-5 and numpy.all(st_p == 0)


def test_one_context_solves_on_two_streams_in_order(otf):
    """Device-pointer solves of one context enqueued without host waits on
    two streams, each batch of a different size (so the shared frame table
    is rewritten between them): every result equals the same solve made
    alone."""
    import torch
    ctx = otf
    p = _lib.default_params()
    dev = torch.device("cuda:0")
    streams = [torch.cuda.current_stream(), torch.cuda.Stream()]
    jobs = []
    for k, shape in enumerate([(96, 130), (150, 64), (64, 64), (120, 200), (80, 90)]):
        f = sim.make_frame(70 + k, shape)
        d = torch.from_numpy(f["data"]).to(dev)
        g = torch.from_numpy(f["gamma"]).to(dev)
        n = _lib.tile_count(*shape)
        jobs.append(dict(f=f, d=d, g=g, out=torch.zeros_like(d),
                         it=torch.zeros(n, dtype=torch.int32, device=dev),
                         st=torch.zeros(n, dtype=torch.int32, device=dev),
                         co=torch.zeros(n, dtype=torch.float32, device=dev), shape=shape))
    torch.cuda.synchronize()

    def enqueue(j, stream):
        ctx.reduce_frames_device([dict(data=j["d"].data_ptr(), gamma=j["g"].data_ptr(),
                                       out=j["out"].data_ptr(), height=j["shape"][0],
                                       width=j["shape"][1], alpha=0.1)], p, None,
                                 j["it"].data_ptr(), j["st"].data_ptr(), j["co"].data_ptr(),
                                 stream.cuda_stream)

    for k, j in enumerate(jobs):
        enqueue(j, streams[k % 2])
    torch.cuda.synchronize()
    for k, j in enumerate(jobs):
        ref = _eager(ctx, [j["f"]["data"]], [j["f"]["gamma"]], [0.1], p)
        assert numpy.array_equal(j["out"].cpu().numpy(), ref[0][0]), k
        assert numpy.array_equal(j["co"].cpu().numpy(), ref[3]), k
        assert numpy.array_equal(j["it"].cpu().numpy(), ref[1]), k
