"""GPU tests of the frame schedule.

A batch's tiles are claimed smallest frame first, while images and every
per-tile array (warm starts in, tile results out) stay in the caller's frame
order.  A tile's result does not depend on which wave solved it or when, so
each frame of a mixed batch must give, bit for bit, what it gives solved
alone.  The batch below is passed large, small, medium, small: its schedule
order differs from the caller's, so a per-tile array indexed by schedule
position instead of caller position fails these tests.
"""
import ctypes

import numpy
import pytest

from ncs_amd import _lib, linstart, masks, sim, warmstart
from ncs_amd.workloads import block_tiles

pytestmark = pytest.mark.gpu

SHAPES = [(150, 203), (40, 36), (96, 80), (33, 47)]  # 130, 9, 30 and 9 blocks
ALPHAS = [0.1, 0.3, 0.05, 0.2]
G = _lib.GEOM_BLOCKS


def _say(*a):
    print(*a, flush=True)


@pytest.fixture
def ctx(gpu_ctx):
    gpu_ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    return gpu_ctx


@pytest.fixture(scope="module")
def frames():
    return [sim.make_frame(600 + k, s) for k, s in enumerate(SHAPES)]


def _counts(idx):
    return [_lib.tile_count(SHAPES[i][0], SHAPES[i][1], G) for i in idx]


def _device_solve(torch, ctx, frames, idx, params, starts=None, noalloc=False):
    """Frames idx (in that order) in one device batch; returns images and
    per-tile iterations, status and cost, in that order."""
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(frames[i]["data"]).to(dev) for i in idx]
    g = [torch.from_numpy(frames[i]["gamma"]).to(dev) for i in idx]
    out = [torch.full_like(x, -1.0) for x in d]
    total = int(sum(_counts(idx)))
    it = torch.full((total,), -7, dtype=torch.int32, device=dev)
    st = torch.full((total,), -7, dtype=torch.int32, device=dev)
    co = torch.zeros(total, dtype=torch.float32, device=dev)
    u0 = None
    if starts is not None:
        u0 = torch.from_numpy(numpy.concatenate([starts[i] for i in idx])).to(dev)
    desc = [dict(data=d[k].data_ptr(), gamma=g[k].data_ptr(), out=out[k].data_ptr(),
                 height=SHAPES[i][0], width=SHAPES[i][1], alpha=ALPHAS[i], geometry=G)
            for k, i in enumerate(idx)]
    stream = torch.cuda.current_stream().cuda_stream
    u0p = None if u0 is None else u0.data_ptr()
    if noalloc:
        lib = ctx.lib
        rc = lib.ncs_hip_reduce_frames_device_noalloc(
            ctx.handle, _lib.frame_array(desc), len(desc), ctypes.byref(params),
            None if u0p is None else ctypes.c_void_p(u0p), ctypes.c_void_p(it.data_ptr()),
            ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(co.data_ptr()),
            ctypes.c_void_p(stream))
        _lib.check(rc, "ncs_hip_reduce_frames_device_noalloc")
    else:
        ctx.reduce_frames_device(desc, params, u0p, it.data_ptr(), st.data_ptr(), co.data_ptr(),
                                 stream)
    torch.cuda.synchronize()
    return ([o.cpu().numpy() for o in out], it.cpu().numpy(), st.cpu().numpy(),
            co.cpu().numpy())


def _split(res, idx):
    """Per-frame pieces of a batch result, keyed by frame number."""
    outs, it, st, co = res
    bounds = numpy.concatenate([[0], numpy.cumsum(_counts(idx))])
    return {i: (outs[k], it[bounds[k]:bounds[k + 1]], st[bounds[k]:bounds[k + 1]],
                co[bounds[k]:bounds[k + 1]]) for k, i in enumerate(idx)}


def _check_against_alone(batch, alone, label):
    for i, ref in alone.items():
        got = batch[i]
        assert numpy.array_equal(got[0], ref[0]), (label, "image", i)
        for k, name in ((1, "iterations"), (2, "status"), (3, "cost")):
            assert numpy.array_equal(got[k], ref[k]), (label, name, i)
        assert numpy.all((got[2] <= 0) & (got[2] >= -5)), (label, i)


def _starts(frames):
    """Explicit per-tile starts: each frame's blocks pulled a quarter of the
    way toward their tile means (padding stays 0)."""
    out = {}
    for k, f in enumerate(frames):
        t, m = block_tiles(f["data"])
        out[k] = (0.75 * t + 0.25 * t.mean(axis=(1, 2), keepdims=True) * m).astype(numpy.float32)
    return out


def _pinned(shape):
    import torch
    t = torch.zeros(shape, dtype=torch.float32, pin_memory=True)
    return t, t.numpy()


@pytest.mark.parametrize("start", ["plain", "linear", "tiles", "learned"])
def test_mixed_batch_matches_each_frame_solved_alone(ctx, frames, start):
    import torch
    p = _lib.default_params()
    starts = None
    if start == "linear":
        ctx.set_linear_start(linstart.load(), linstart.load_slope())
        p = _lib.default_params(start=_lib.START_LINEAR)
    elif start == "tiles":
        starts = _starts(frames)
        p = _lib.default_params(start=_lib.START_TILES)
    elif start == "learned":
        warmstart.set_context_weights(warmstart.load())
        p = _lib.default_params(start=_lib.START_LEARNED)
    idx = list(range(len(SHAPES)))
    _say("start", start, "batch order", idx, "tiles", _counts(idx))
    batch = _split(_device_solve(torch, ctx, frames, idx, p, starts), idx)
    alone = {i: _split(_device_solve(torch, ctx, frames, [i], p, starts), [i])[i] for i in idx}
    _check_against_alone(batch, alone, start)
    # and the reverse caller order, whose schedule is another permutation
    rev = idx[::-1]
    _check_against_alone(_split(_device_solve(torch, ctx, frames, rev, p, starts), rev), alone,
                         start + " reversed")


def test_noalloc_and_host_batch_keep_the_callers_order(ctx, frames):
    import torch
    ctx.set_linear_start(linstart.load(), linstart.load_slope())
    p = _lib.default_params(start=_lib.START_LINEAR)
    idx = list(range(len(SHAPES)))
    alone = {i: _split(_device_solve(torch, ctx, frames, [i], p), [i])[i] for i in idx}
    _say("allocation-free entry point")
    _check_against_alone(_split(_device_solve(torch, ctx, frames, idx, p, noalloc=True), idx),
                         alone, "noalloc")
    _say("host-memory batch")
    outs = [numpy.zeros(s, numpy.float32) for s in SHAPES]
    it, st, co = ctx.reduce_frames([f["data"] for f in frames], [f["gamma"] for f in frames],
                                   outs, ALPHAS, G, p)
    _check_against_alone(_split((outs, it, st, co), idx), alone, "host")


def test_pipeline_and_net_predictions_keep_the_callers_order(ctx, frames):
    import torch
    p = _lib.default_params()
    idx = list(range(len(SHAPES)))
    alone = {i: _split(_device_solve(torch, ctx, frames, [i], p), [i])[i] for i in idx}
    _say("pipeline")
    pl = _lib.Pipeline(ctx, SHAPES, ALPHAS, params=p, geometry=G, keep_tile_results=True)
    try:
        for i, f in enumerate(frames):
            pl.set_gamma(i, f["gamma"])
        ins, outs, keep = [], [], []
        for f in frames:
            ti, a = _pinned(f["data"].shape)
            to, b = _pinned(f["data"].shape)
            a[...] = f["data"]
            keep += [ti, to]
            ins.append(a)
            outs.append(b)
        pl.launch(ins, outs)
        pl.wait()
        outs = [b.copy() for b in outs]
        it, st, co = pl.tile_results()
    finally:
        pl.close()
    _check_against_alone(_split((outs, it, st, co), idx), alone, "pipeline")
    _say("net predictions")
    warmstart.set_context_weights(warmstart.load())
    dev = torch.device("cuda:0")

    def predict(order):
        d = [torch.from_numpy(frames[i]["data"]).to(dev) for i in order]
        g = [torch.from_numpy(frames[i]["gamma"]).to(dev) for i in order]
        total = int(sum(_counts(order)))
        u0 = torch.full((total, 256), -1.0, dtype=torch.float32, device=dev)
        desc = [dict(data=d[k].data_ptr(), gamma=g[k].data_ptr(), out=None,
                     height=SHAPES[i][0], width=SHAPES[i][1], alpha=ALPHAS[i], geometry=G)
                for k, i in enumerate(order)]
        ctx.predict_frames_device(desc, u0.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return u0.cpu().numpy()

    mixed = predict(idx)
    bounds = numpy.concatenate([[0], numpy.cumsum(_counts(idx))])
    for k, i in enumerate(idx):
        one = predict([i])
        assert numpy.array_equal(mixed[bounds[k]:bounds[k + 1]], one), ("net", i)
