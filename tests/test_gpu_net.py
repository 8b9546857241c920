"""GPU checks of the warm-start conv net kernels (BF16 MFMA) against the FP64
CPU oracle (oracle/convnet_ref.py, with bf16 rounding of the stored
operands), and of the learned start inside the solver."""
import numpy
import pytest

import oracle
from oracle import convnet_ref as ref
from ncs_amd import _lib, masks, sim, warmstart

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(autouse=True)
def _needs_gpu(gpu_ctx):
    return gpu_ctx


def _inputs(n, seed):
    import torch
    rng = numpy.random.default_rng(seed)
    fr = sim.make_frame(seed, (64, 64))
    d = warmstart.reference_windows(fr["data"])[:n].astype(numpy.float32)
    g = warmstart.reference_windows(fr["gamma"])[:n].astype(numpy.float32)
    del rng
    td, tg = torch.from_numpy(d).cuda(), torch.from_numpy(g).cuda()
    x, sigma = warmstart.normalize(td, tg)
    return d, g, x, sigma


def test_forward_matches_oracle():
    import torch
    _say("net forward vs oracle")
    p = warmstart.init_params(7, last_scale=1.0)
    d, g, x, _ = _inputs(12, 3)
    y = warmstart.NetFunction.apply(torch.from_numpy(p).cuda(), x)
    torch.cuda.synchronize()
    yr, _ = ref.forward(p, x.cpu().numpy().astype(numpy.float64), bf16=True)
    err = numpy.max(numpy.abs(y.cpu().numpy() - yr))
    scale = numpy.max(numpy.abs(yr))
    _say("forward max abs err", err, "scale", scale)
    # the oracle rounds the same operands to BF16; what remains is FP32 vs
    # FP64 accumulation (measured 1.7e-6 of the scale) and the odd BF16
    # rounding-boundary flip of an activation it causes
    assert err <= 1e-4 * scale
    # the same net without BF16 rounding: what is left is the rounding itself
    yf, _ = ref.forward(p, x.cpu().numpy().astype(numpy.float64), bf16=False)
    errf = numpy.max(numpy.abs(y.cpu().numpy() - yf))
    _say("forward vs unrounded oracle max abs err", errf)
    assert errf <= 5e-2 * scale


def test_backward_matches_oracle():
    import torch
    _say("net backward vs oracle")
    p = warmstart.init_params(8, last_scale=1.0)
    d, g, x, _ = _inputs(9, 4)
    tp = torch.from_numpy(p).cuda().requires_grad_(True)
    y = warmstart.NetFunction.apply(tp, x)
    gy = torch.from_numpy(numpy.random.default_rng(5).standard_normal((9, 16, 16)).astype(numpy.float32)).cuda()
    (y * gy).sum().backward()
    torch.cuda.synchronize()
    gk = tp.grad.cpu().numpy().astype(numpy.float64)
    yr, cache = ref.forward(p, x.cpu().numpy().astype(numpy.float64), bf16=True)
    gr = ref.backward(p, cache, gy.cpu().numpy().astype(numpy.float64), bf16=True)
    for li, (ws, bs, cin, cout) in enumerate(warmstart.layer_slices()):
        for name, sl in (("w", ws), ("b", bs)):
            num = numpy.linalg.norm(gk[sl] - gr[sl])
            den = numpy.linalg.norm(gr[sl])
            _say("layer", li + 1, name, "rel err", num / den)
            # measured at most 5.7e-3 (layer 1, whose gradient sums over
            # the most BF16-rounded products)
            assert num <= 1.5e-2 * den, (li, name)


def test_learned_start_matches_plain_cost_and_tensor_path():
    """Fused gather + net + solve (NCS_START_LEARNED) reaches the plain
    solver's per-tile costs; the net's start equals the tensor path's."""
    import torch
    _say("learned start end to end")
    mask = masks.otf_mask()
    mf = masks.to_fft_order(mask).astype(numpy.float32)
    ctx = warmstart.set_context_weights(warmstart.init_params(9))
    ctx.set_otf_mask(mf)
    fr = sim.make_frame(21, (96, 112))
    td = torch.from_numpy(fr["data"]).cuda()
    tg = torch.from_numpy(fr["gamma"]).cuda()
    nt = _lib.tile_count(96, 112)
    outs = {}
    for start in (_lib.START_PLAIN, _lib.START_LEARNED):
        out = torch.zeros_like(td)
        it = torch.zeros(nt, dtype=torch.int32, device="cuda")
        st = torch.zeros_like(it)
        cs = torch.zeros(nt, device="cuda")
        ctx.reduce_frames_device([dict(data=td.data_ptr(), gamma=tg.data_ptr(), out=out.data_ptr(),
                                       height=96, width=112, alpha=0.1)],
                                 _lib.default_params(start=start), None, it.data_ptr(),
                                 st.data_ptr(), cs.data_ptr(),
                                 torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        outs[start] = (out.cpu().numpy(), it.cpu().numpy(), st.cpu().numpy(), cs.cpu().numpy())
    pc, wc = outs[_lib.START_PLAIN][3], outs[_lib.START_LEARNED][3]
    rel = numpy.abs(wc - pc) / numpy.abs(pc)
    _say("learned vs plain: cost max rel", rel.max(), "iters", outs[0][1].mean(), outs[2][1].mean())
    assert numpy.all(outs[_lib.START_LEARNED][2] == 0)
    assert rel.max() <= 1e-5
    # tensor path: normalise + NetFunction + tile start gives the same costs
    d = warmstart.reference_windows(fr["data"]).astype(numpy.float32)
    g = warmstart.reference_windows(fr["gamma"]).astype(numpy.float32)
    tdd, tgg = torch.from_numpy(d).cuda(), torch.from_numpy(g).cuda()
    u0 = warmstart.predict_u0(torch.from_numpy(warmstart.init_params(9)).cuda(), tdd, tgg).detach()
    u, it2, st2, c2 = ctx.solve_tiles(d, g, 0.1, _lib.default_params())
    outd = torch.zeros_like(tdd)
    it3 = torch.zeros(len(d), dtype=torch.int32, device="cuda")
    st3 = torch.zeros_like(it3)
    c3 = torch.zeros(len(d), device="cuda")
    ctx.solve_tiles_device(tdd.data_ptr(), tgg.data_ptr(), u0.contiguous().data_ptr(),
                           outd.data_ptr(), it3.data_ptr(), st3.data_ptr(), c3.data_ptr(), len(d),
                           0.1, _lib.default_params(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rel2 = numpy.abs(c3.cpu().numpy() - wc) / numpy.abs(wc)
    assert rel2.max() <= 1e-5


def test_training_lowers_the_loss_and_saves_iterations():
    """A short training run on solver-made pairs lowers the loss; the trained
    start needs fewer iterations than the plain start at matched cost."""
    import torch
    _say("short training run")
    d, g, u = warmstart.make_pairs(4096, seed0=20_000_000, frame_shape=(128, 128), log=_say)
    params, hist = warmstart.train(d, g, u, steps=300, batch=512, lr=2e-3, log=_say)
    assert hist[-1][1] < 0.7 * hist[0][1]
    dt, gt_, ut = warmstart.make_pairs(1024, seed0=1, frame_shape=(128, 128), log=_say)
    u0 = warmstart.predict_u0(params, dt, gt_).detach()
    ctx = _lib.context(0)
    n = dt.shape[0]
    for precond in (0, 1):
        res = {}
        for name, start in (("plain", None), ("learned", u0)):
            out = torch.zeros_like(dt)
            it = torch.zeros(n, dtype=torch.int32, device="cuda")
            st = torch.zeros_like(it)
            cs = torch.zeros(n, device="cuda")
            ctx.solve_tiles_device(dt.data_ptr(), gt_.data_ptr(),
                                   None if start is None else start.contiguous().data_ptr(),
                                   out.data_ptr(), it.data_ptr(), st.data_ptr(), cs.data_ptr(), n,
                                   0.1, _lib.default_params(precond=precond),
                                   torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            res[name] = (it.float().mean().item(), cs.cpu().numpy())
        rel = numpy.abs(res["learned"][1] - res["plain"][1]) / numpy.abs(res["plain"][1])
        _say("precond %d: iterations plain %.2f learned %.2f, cost max rel %.2e"
             % (precond, res["plain"][0], res["learned"][0], rel.max()))
        assert rel.max() <= 1e-5
        if precond == 0:  # the reference's L-BFGS (scalar initial Hessian)
            assert res["learned"][0] < res["plain"][0]
