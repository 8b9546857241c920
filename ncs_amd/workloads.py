"""Synthetic inputs of the benchmark workloads, shared by bench.py and the
GPU tests: the uneven batch (mixed frame sizes and ROIs, heavy-tailed
variance, alpha per frame), the raw-ADU time series, and the zero-padded
block view of a frame that the block tiling solves.
"""
import numpy


def uneven_frame(args):
    from . import sim
    seed, shape = args
    fr = sim.make_frame(seed, shape, heavy_tail=True)
    return fr["data"], fr["gamma"]


def uneven_specs(seed=3000):
    rng = numpy.random.default_rng(seed)
    shapes = [(512, 512)] * 8 + [(1024, 2048)] * 4 + [(2304, 2304)] * 2
    shapes += [tuple(int(v) for v in rng.integers(64, 401, 2)) for _ in range(16)]
    alphas = [float(a) for a in rng.uniform(0.05, 0.3, len(shapes))]
    seeds = [seed + 1 + i for i in range(len(shapes))]
    return shapes, alphas, seeds


def block_tiles(img, fill=0.0):
    """Zero-padded 16 x 16 blocks of a frame, raster order, and their 0/1
    validity masks (the block tiling of the solver)."""
    H, W = img.shape
    th, tw = (H + 15) // 16, (W + 15) // 16
    p = numpy.full((th * 16, tw * 16), fill, numpy.float32)
    p[:H, :W] = img
    v = numpy.zeros((th * 16, tw * 16), numpy.float32)
    v[:H, :W] = 1.0
    t = p.reshape(th, 16, tw, 16).transpose(0, 2, 1, 3).reshape(-1, 16, 16)
    m = v.reshape(th, 16, tw, 16).transpose(0, 2, 1, 3).reshape(-1, 16, 16)
    return t, m


def stream_series(n, shape, seed, dev, blink_p=0.3, segment_um=1.0):
    """Raw uint16 ADU frames of a blinking filament scene (pinned host
    memory), rendered on the GPU with torch: each 1 um filament segment is on
    with probability blink_p per frame, background drifts linearly from 10 to
    40 photons/pixel, camera maps fixed.  Returns (adu [n, H, W] uint16
    numpy view, its pinned torch owner, offset, gain, var)."""
    import torch
    from . import sim
    rng = numpy.random.default_rng(seed)
    H, W = shape
    n_fil = max(1, int(round(sim.FILAMENTS_PER_UM2 * H * W * sim.PIXEL_UM ** 2)))
    fils = sim.wlc_filaments(rng, n_fil, H * sim.PIXEL_UM, W * sim.PIXEL_UM)
    seg = max(1, int(round(segment_um / 0.01)))
    seg_id, nseg = [], 0
    for f in fils:
        k = (len(f) + seg - 1) // seg
        seg_id.append(nseg + numpy.arange(len(f)) // seg)
        nseg += k
    pts = torch.from_numpy(numpy.concatenate(fils, axis=0) / sim.PIXEL_UM).to(dev)
    sid = torch.from_numpy(numpy.concatenate(seg_id)).to(dev)
    offset, gain, var = sim.camera_maps(rng, shape)
    t_off = torch.from_numpy(offset).to(dev)
    t_gain = torch.from_numpy(gain).to(dev)
    t_sd = torch.from_numpy(numpy.sqrt(var)).to(dev)
    k = 1.0 / (numpy.sqrt(2.0) * sim.SIGMA_PX)
    offs = torch.arange(-3, 4, device=dev, dtype=torch.float64)
    base = torch.floor(pts)
    ix = base[:, 0:1] + offs
    iy = base[:, 1:2] + offs
    ex = 0.5 * (torch.erf((ix + 1 - pts[:, 0:1]) * k) - torch.erf((ix - pts[:, 0:1]) * k))
    ey = 0.5 * (torch.erf((iy + 1 - pts[:, 1:2]) * k) - torch.erf((iy - pts[:, 1:2]) * k))
    ok = ((ix >= 0) & (ix < H))[:, :, None] & ((iy >= 0) & (iy < W))[:, None, :]
    idx = (ix[:, :, None] * W + iy[:, None, :]).long()
    val = (ex[:, :, None] * ey[:, None, :]) * ok
    idx = torch.where(ok, idx, torch.zeros_like(idx)).view(len(pts), -1)
    val = val.view(len(pts), -1)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    owner = torch.empty((n, H, W), dtype=torch.int16, pin_memory=True)
    adu_np = owner.numpy().view(numpy.uint16)
    per_point = 2000.0 * 0.01
    for t in range(n):
        on = (torch.rand(nseg, generator=gen, device=dev) < blink_p).double()
        w = per_point * on[sid]
        img = torch.zeros(H * W, dtype=torch.float64, device=dev)
        img.index_add_(0, idx.reshape(-1), (val * w[:, None]).reshape(-1))
        bg = 10.0 + 30.0 * t / max(1, n - 1)
        ideal = img.view(H, W) + bg
        ph = torch.poisson(ideal, generator=gen)
        a = t_gain * ph + t_sd * torch.randn(ideal.shape, generator=gen, device=dev,
                                             dtype=torch.float64) + t_off
        a = torch.clamp(torch.round(a), 0, 65535).to(torch.int32)
        adu_np[t] = a.cpu().numpy().astype(numpy.uint16)
    return adu_np, owner, offset, gain, var
