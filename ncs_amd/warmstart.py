"""Learned warm start for the per-tile L-BFGS: a 4-layer 3x3 conv net.

The net (2 -> 32 -> 32 -> 32 -> 1 channels, ReLU after layers 1-3, zero
padding inside each 16 x 16 tile) runs on hand-written HIP kernels
(ncs_convnet.hip: BF16 MFMA, FP32 accumulation) for inference and for
training; ``NetFunction`` exposes them to torch.autograd so the optimiser is
ordinary torch plumbing.

Per tile, with d the photo-electron data and gamma the read-noise map:
    mu = mean(d),  sigma = sqrt(max(mu, 1) + mean(gamma)),
    x = [(d - mu) / sigma, gamma / sigma^2],   u0 = d + sigma * net(x).
The net is trained on (tile, NCS solution) pairs produced by the plain GPU
solver itself; the solver keeps the plain start u0 = d for every tile whose
data scores a lower NCS cost than the net's start (per-tile fallback).

No counterpart in the reference: its solvers always start from the data
(clib/pyCNCS/ncs.c:549-551, opencl/ncs.cl:906).
"""
import ctypes
import math
import os
import time

import numpy

from . import _lib, masks, sim

N_PARAMS = _lib.NET_PARAMS
LAYERS = [(2, 32), (32, 32), (32, 32), (32, 1)]  # (cin, cout)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "warmstart_otf_a0.1.npy")


def layer_slices():
    """(weight slice, bias slice, cin, cout) of each layer in the flat vector."""
    out, off = [], 0
    for cin, cout in LAYERS:
        nw = cout * cin * 9
        out.append((slice(off, off + nw), slice(off + nw, off + nw + cout), cin, cout))
        off += nw + cout
    assert off == N_PARAMS
    return out


def init_params(seed=0, last_scale=0.1):
    """He-uniform weights, zero biases; the last layer is scaled down so the
    untrained net starts close to the plain start."""
    rng = numpy.random.default_rng(seed)
    p = numpy.zeros(N_PARAMS, numpy.float32)
    for li, (ws, bs, cin, cout) in enumerate(layer_slices()):
        bound = math.sqrt(6.0 / (cin * 9))
        w = rng.uniform(-bound, bound, ws.stop - ws.start)
        if li == len(LAYERS) - 1:
            w *= last_scale
        p[ws] = w
    return p


def normalize(d, g):
    """Net inputs [n, 2, 16, 16] and sigma [n, 1, 1] of tiles d, g [n, 16, 16]
    (torch tensors).  Must match tile_inputs() in ncs_convnet.hip."""
    import torch
    mu = d.mean(dim=(-2, -1), keepdim=True)
    sigma = torch.sqrt(torch.clamp(mu, min=1.0) + g.mean(dim=(-2, -1), keepdim=True))
    x = torch.stack([(d - mu) / sigma, g / (sigma * sigma)], dim=1)
    return x.contiguous(), sigma


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class NetFunction(object):
    """torch.autograd.Function over ncs_hip_net_forward / _backward."""
    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is None:
            import torch

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, params, x):
                    if not (params.is_cuda and x.is_cuda):
                        raise _lib.NCSHipError("the warm-start net runs on the GPU only")
                    params = params.contiguous()
                    x = x.contiguous()
                    n = x.shape[0]
                    y = torch.empty((n, 16, 16), dtype=torch.float32, device=x.device)
                    acts = torch.empty((3, n, 256, 32), dtype=torch.bfloat16, device=x.device)
                    c = _lib.context(x.device.index or 0)
                    _lib.check(c.lib.ncs_hip_net_forward(c.handle, params.data_ptr(), x.data_ptr(),
                                                         y.data_ptr(), acts.data_ptr(), n,
                                                         _stream()),
                               "ncs_hip_net_forward")
                    ctx.save_for_backward(params, x, acts)
                    return y

                @staticmethod
                def backward(ctx, gy):
                    params, x, acts = ctx.saved_tensors
                    gy = gy.contiguous().float()
                    gp = torch.empty_like(params)
                    c = _lib.context(x.device.index or 0)
                    _lib.check(c.lib.ncs_hip_net_backward(c.handle, params.data_ptr(),
                                                          x.data_ptr(), acts.data_ptr(),
                                                          gy.data_ptr(), gp.data_ptr(), None,
                                                          x.shape[0], _stream()),
                               "ncs_hip_net_backward")
                    return gp, None

            cls._cls = _Fn
        return cls._cls

    @classmethod
    def apply(cls, params, x):
        return cls.get().apply(params, x)


def predict_u0(params, d, g):
    """Warm start u0 of tiles d, g [n, 16, 16] (device tensors)."""
    x, sigma = normalize(d, g)
    return d + sigma * NetFunction.apply(params, x)


# ---------------------------------------------------------------------------
# training pairs
# ---------------------------------------------------------------------------

def reference_windows(img, ref_geometry=True):
    """All 16 x 16 windows of the reference tiling (ncs.c:71-113) of a frame,
    [tiles, 16, 16], raster order (host helper for training-set assembly)."""
    H, W = img.shape
    p = numpy.pad(img, 1, "edge")

    def starts(n):
        out, i = [], -1
        while True:
            b = n + 1 - 16 if i + 16 > n + 1 else i
            out.append(b + 1)
            if b == n + 1 - 16:
                return out
            i += 14

    ys, xs = starts(H), starts(W)
    return numpy.stack([p[y:y + 16, x:x + 16] for y in ys for x in xs])


def _frame_windows(args):
    seed, shape = args
    fr = sim.make_frame(seed, shape)
    return reference_windows(fr["data"]), reference_windows(fr["gamma"])


def gen_windows(n_tiles, seed0=10_000_000, frame_shape=(512, 512), workers=8, log=print):
    """Host windows (d, gamma) [n, 16, 16] of synthetic frames with seeds
    >= seed0.  Uses a process pool: call it before the GPU is initialised."""
    per = _lib.tile_count(*frame_shape)
    nfr = (n_tiles + per - 1) // per
    t0 = time.time()
    jobs = [(seed0 + i, frame_shape) for i in range(nfr)]
    if workers > 1:
        from multiprocessing import get_context
        with get_context("fork").Pool(workers) as pool:
            res = pool.map(_frame_windows, jobs)
    else:
        res = [_frame_windows(j) for j in jobs]
    d = numpy.concatenate([r[0] for r in res])[:n_tiles].astype(numpy.float32)
    g = numpy.concatenate([r[1] for r in res])[:n_tiles].astype(numpy.float32)
    log("windows: %d tiles from %d frames in %.1f s" % (len(d), nfr, time.time() - t0))
    return d, g


def solve_pairs(d, g, alpha=0.1, mask=None, params=None, device=0, log=print):
    """Device tensors (d, gamma, u*) with u* from the plain GPU solver."""
    import torch
    ctx = _lib.context(device)
    m = mask if mask is not None else masks.otf_mask()
    ctx.set_otf_mask(masks.to_fft_order(m).astype(numpy.float32))
    dev = torch.device("cuda", device)
    td, tg = torch.from_numpy(d).to(dev), torch.from_numpy(g).to(dev)
    u = torch.empty_like(td)
    it = torch.empty(len(d), dtype=torch.int32, device=dev)
    st = torch.empty_like(it)
    p = params if params is not None else _lib.default_params()
    ctx.solve_tiles_device(td.data_ptr(), tg.data_ptr(), None, u.data_ptr(), it.data_ptr(),
                           st.data_ptr(), None, len(d), alpha, p, _stream())
    torch.cuda.synchronize()
    log("pairs solved: iterations mean %.2f" % it.float().mean().item())
    return td, tg, u


def make_pairs(n_tiles, seed0=10_000_000, frame_shape=(512, 512), alpha=0.1, mask=None,
               params=None, device=0, workers=1, log=print):
    """(d, gamma, u*) device tensors: gen_windows + solve_pairs."""
    d, g = gen_windows(n_tiles, seed0, frame_shape, workers, log)
    return solve_pairs(d, g, alpha, mask, params, device, log)


def train(d, g, u, steps=3000, batch=1024, lr=2e-3, seed=0, log=print, params0=None):
    """Adam on the MSE of the normalised correction (u* - d) / sigma; returns
    the trained FP32 parameters (device tensor) and the loss history."""
    import torch
    dev = d.device
    p = torch.nn.Parameter(torch.from_numpy(params0 if params0 is not None
                                            else init_params(seed)).to(dev))
    opt = torch.optim.Adam([p], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    x_all, sig_all = normalize(d, g)
    target = (u - d) / sig_all
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    hist = []
    t0 = time.time()
    for s in range(steps):
        idx = torch.randint(0, d.shape[0], (batch,), device=dev, generator=gen)
        y = NetFunction.apply(p, x_all[idx])
        loss = torch.mean((y - target[idx]) ** 2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if s % 200 == 0 or s == steps - 1:
            hist.append((s, float(loss.item())))
            log("train step %d loss %.5f (%.1f s)" % (s, hist[-1][1], time.time() - t0))
    return p.detach(), hist


def save(params, path=DEFAULT_WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    numpy.save(path, numpy.asarray(params, dtype=numpy.float32))


def load(path=DEFAULT_WEIGHTS):
    return numpy.load(path, allow_pickle=False).astype(numpy.float32)


def set_context_weights(params, device=0):
    """Installs weights for NCS_START_LEARNED on the device's context."""
    ctx = _lib.context(device)
    p = numpy.ascontiguousarray(params, dtype=numpy.float32)
    if p.shape != (N_PARAMS,):
        raise ValueError("expected %d parameters" % N_PARAMS)
    _lib.check(ctx.lib.ncs_hip_net_set_weights(ctx.handle, ctypes.c_void_p(p.ctypes.data)),
               "ncs_hip_net_set_weights")
    return ctx


def main(argv=None):
    """python -m ncs_amd.warmstart train [--tiles N] [--steps S] [--out PATH]

    Generates (tile, NCS solution) pairs with the plain GPU solver from
    synthetic frames (seeds >= 1e7, disjoint from the benchmark's seeds
    1..10), trains the net with its HIP kernels and saves the FP32 weights."""
    import argparse
    import json

    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--tiles", type=int, default=200_000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--alpha", type=float, default=0.1)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, flush=True)
    # every process pool runs before the GPU is touched
    dn, gn = gen_windows(a.tiles, workers=a.workers, log=log)
    dh, gh = gen_windows(8192, seed0=1, workers=a.workers, log=log)
    d, g, u = solve_pairs(dn, gn, alpha=a.alpha, log=log)
    params, hist = train(d, g, u, steps=a.steps, batch=a.batch, lr=a.lr, log=log)
    save(params.cpu().numpy(), a.out)
    # held-out tiles (seeds 1..): iterations of the plain and the learned start
    dt, gt_, ut = solve_pairs(dh, gh, alpha=a.alpha, log=log)
    u0 = predict_u0(params, dt, gt_).detach().contiguous()
    ctx = _lib.context(0)
    n = dt.shape[0]
    res = {}
    for name, start in (("plain", None), ("learned", u0)):
        out = torch.zeros_like(dt)
        it = torch.zeros(n, dtype=torch.int32, device=dt.device)
        st = torch.zeros_like(it)
        cs = torch.zeros(n, device=dt.device)
        ctx.solve_tiles_device(dt.data_ptr(), gt_.data_ptr(),
                               None if start is None else start.data_ptr(), out.data_ptr(),
                               it.data_ptr(), st.data_ptr(), cs.data_ptr(), n, a.alpha,
                               _lib.default_params(), _stream())
        torch.cuda.synchronize()
        res[name] = dict(iterations=float(it.float().mean().item()), cost=cs.cpu().numpy())
    rel = numpy.abs(res["learned"]["cost"] - res["plain"]["cost"]) / numpy.abs(res["plain"]["cost"])
    err_net = float(((u0 - ut).abs().mean() / ut.abs().mean()).item())
    err_plain = float(((dt - ut).abs().mean() / ut.abs().mean()).item())
    summary = dict(weights=a.out, tiles=a.tiles, steps=a.steps, loss=hist,
                   held_out_iterations_plain=res["plain"]["iterations"],
                   held_out_iterations_learned=res["learned"]["iterations"],
                   held_out_cost_max_rel=float(rel.max()),
                   held_out_rel_error_start_net=err_net, held_out_rel_error_start_plain=err_plain)
    log(json.dumps(summary))
    return summary


if __name__ == "__main__":
    main()
