"""GPU check of the allocation-free entry point
(ncs_hip_reduce_frames_device_noalloc), through the C ABI, against
ncs_hip_reduce_frames_device on the same inputs: images, iterations, status
and per-tile costs bit for bit, on two calls in a row, the second with a
smaller batch (fewer frames, fewer tiles), and its refusals."""
import ctypes

import numpy
import pytest

from ncs_amd import _lib, linstart, masks, sim

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _batch(torch, shapes, seed, geometry):
    dev = torch.device("cuda:0")
    frames, keep = [], []
    for k, shape in enumerate(shapes):
        f = sim.make_frame(seed + k, shape)
        d = torch.from_numpy(f["data"]).to(dev)
        g = torch.from_numpy(f["gamma"]).to(dev)
        keep += [d, g]
        frames.append(dict(data=d, gamma=g, height=shape[0], width=shape[1],
                           alpha=0.05 + 0.05 * k, geometry=geometry))
    total = sum(_lib.tile_count(h, w, geometry) for h, w in shapes)
    return frames, total, keep


def _solve(torch, lib, ctx, frames, total, params, noalloc):
    dev = torch.device("cuda:0")
    outs = [torch.full_like(f["data"], -1.0) for f in frames]
    it = torch.full((total,), -7, dtype=torch.int32, device=dev)
    st = torch.full((total,), -7, dtype=torch.int32, device=dev)
    co = torch.zeros(total, dtype=torch.float32, device=dev)
    arr = _lib.frame_array([dict(data=f["data"].data_ptr(), gamma=f["gamma"].data_ptr(),
                                 out=outs[i].data_ptr(), height=f["height"], width=f["width"],
                                 alpha=f["alpha"], geometry=f["geometry"])
                            for i, f in enumerate(frames)])
    fn = (lib.ncs_hip_reduce_frames_device_noalloc if noalloc
          else lib.ncs_hip_reduce_frames_device)
    rc = fn(ctx.handle, arr, len(frames), ctypes.byref(params), None, ctypes.c_void_p(it.data_ptr()),
            ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(co.data_ptr()),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return ([o.cpu().numpy() for o in outs], it.cpu().numpy(), st.cpu().numpy(),
            co.cpu().numpy())


def _same(a, b):
    assert len(a[0]) == len(b[0])
    for x, y in zip(a[0], b[0]):
        assert numpy.array_equal(x, y)
    for k in (1, 2, 3):
        assert numpy.array_equal(a[k], b[k])


@pytest.mark.parametrize("start", ["plain", "linear"])
def test_noalloc_matches_the_allocating_entry_point(gpu_ctx, start):
    import torch
    lib = _lib.load()
    ctx = gpu_ctx
    ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    ctx.set_linear_start(linstart.load())
    p = _lib.default_params(start=_lib.START_LINEAR if start == "linear" else _lib.START_PLAIN)
    # 30 frames (two kernel-argument chunks of the frame table), then 2
    big, nbig, keep1 = _batch(torch, [(64 + 8 * (k % 5), 90 + 6 * (k % 7)) for k in range(30)],
                              500, _lib.GEOM_REFERENCE)
    small, nsmall, keep2 = _batch(torch, [(48, 70), (33, 40)], 600, _lib.GEOM_BLOCKS)
    batches = ((big, nbig, "30 frames"), (small, nsmall, "2 frames"))
    refs = [_solve(torch, lib, ctx, f, n, p, noalloc=False) for f, n, _ in batches]
    # two calls in a row, the second smaller
    gots = [_solve(torch, lib, ctx, f, n, p, noalloc=True) for f, n, _ in batches]
    for (f, n, label), ref, got in zip(batches, refs, gots):
        _say(start, label, n, "tiles: status", numpy.unique(got[2]).tolist())
        assert numpy.all(got[2] == 0)
        _same(got, ref)


def test_noalloc_refusals(gpu_ctx):
    import torch
    lib = _lib.load()
    ctx = gpu_ctx
    ctx.set_otf_mask(masks.to_fft_order(masks.otf_mask()).astype(numpy.float32))
    frames, total, keep = _batch(torch, [(32, 32)], 700, _lib.GEOM_REFERENCE)
    arr = _lib.frame_array([dict(data=frames[0]["data"].data_ptr(),
                                 gamma=frames[0]["gamma"].data_ptr(), out=None, height=32,
                                 width=32, alpha=0.1)] * 257)
    rc = lib.ncs_hip_reduce_frames_device_noalloc(ctx.handle, arr, 257,
                                                  ctypes.byref(_lib.default_params()), None,
                                                  None, None, None, None)
    assert rc == -1 and "256" in _lib.last_error()
    rc = lib.ncs_hip_reduce_frames_device_noalloc(
        ctx.handle, arr, 1, ctypes.byref(_lib.default_params(start=_lib.START_LEARNED)), None,
        None, None, None, None)
    assert rc == -1 and "learned" in _lib.last_error()
