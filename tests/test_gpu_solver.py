"""GPU parity of the per-tile evaluation and solver against the FP64 oracle.

Tolerances: cost/gradient evaluation as the reference's own tests
(opencl/pyOpenCLNCS/test/test_calc_nc.py:186,203: max relative gradient
error with floor 1 < 1e-5); solver per-tile final cost within 1e-5
relative of the oracle's (|f| as denominator), the bar of the build's
contract, and max |du| / max(u) < 1e-2 (test_solver.py:101-102).
"""
import numpy
import pytest

import oracle
from ncs_amd import _lib, masks, sim

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _tiles_from_frame(frame, n, rng):
    d, g = frame["data"], frame["gamma"]
    H, W = d.shape
    ys = rng.integers(0, H - 16, n)
    xs = rng.integers(0, W - 16, n)
    return (numpy.stack([d[y:y + 16, x:x + 16] for y, x in zip(ys, xs)]),
            numpy.stack([g[y:y + 16, x:x + 16] for y, x in zip(ys, xs)]))


@pytest.mark.parametrize("mask", ["otf", "gauss", "random"])
def test_eval_matches_golden_fp64(gpu_ctx, golden, mask):
    _say("eval parity, mask", mask)
    mf = golden["mask_" + mask].astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    d = golden["eval_d"].astype(numpy.float32)
    g = golden["eval_g"].astype(numpy.float32)
    u = golden["eval_u"].astype(numpy.float32)
    cost, grad = gpu_ctx.eval_tiles(d, g, u, 0.1)
    # reference values at the float32-rounded inputs, in FP64
    for i in range(d.shape[0]):
        fr, gr = oracle.cost_grad(u[i].astype(numpy.float64), d[i].astype(numpy.float64),
                                  g[i].astype(numpy.float64), mf.astype(numpy.float64), 0.1)
        assert abs(cost[i] - fr) <= 2e-6 * abs(fr), (i, cost[i], fr)
        rel = numpy.abs(grad[i] - gr) / numpy.maximum(1.0, numpy.abs(gr))
        assert rel.max() < 1e-5, (i, rel.max())


@pytest.mark.parametrize("mask", ["gauss", "random"])
def test_solver_reference_test_case(gpu_ctx, mask):
    """opencl/pyOpenCLNCS/test/test_solver.py:104-163 inputs (seed 1, ten
    U(10,20) tiles, U(2,4) gamma, alpha 0.1) against the oracle."""
    _say("solver ref case, mask", mask)
    rng = numpy.random.RandomState(1)
    data = rng.uniform(10.0, 20.0, (10, 16, 16)).astype(numpy.float32)
    gamma = rng.uniform(2.0, 4.0, (10, 16, 16)).astype(numpy.float32)
    m = masks.gaussian_highpass_mask() if mask == "gauss" else masks.random_mask(16, numpy.random.default_rng(4))
    mf = masks.to_fft_order(m).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    u, it, st, cost = gpu_ctx.solve_tiles(data, gamma, 0.1, _lib.default_params())
    ref = oracle.solve_tiles(data, gamma, mf, 0.1, f32=True)
    _say("iters gpu", it.tolist(), "oracle", ref["iters"].tolist(), "status", st.tolist())
    assert numpy.all(st == 0)
    f64, _, _ = oracle.eval_cost_f32(u, data, gamma, mf, 0.1)
    rel = numpy.abs(f64 - ref["cost"]) / numpy.abs(ref["cost"])
    assert rel.max() <= 1e-5, rel.max()
    assert numpy.all(numpy.abs(cost - f64) <= 1e-5 * numpy.abs(f64))
    for i in range(10):
        nd = numpy.max(numpy.abs(u[i] - ref["u"][i])) / numpy.max(ref["u"][i])
        assert nd < 1e-2, (i, nd)


@pytest.mark.parametrize("precond", [0, 1, 2])
@pytest.mark.parametrize("history", [1, 3, 5, 6, 8])
def test_solver_synthetic_tiles_every_history(gpu_ctx, history, precond):
    """Every initial inverse Hessian (liblbfgs's scalar s.y/y.y, the spectral
    one and the curvature-corrected one) reaches the oracle's minimiser for
    every history length."""
    _say("synthetic tiles, history", history, "precond", precond)
    fr = sim.make_frame(7, (128, 128))
    rng = numpy.random.default_rng(8)
    d, g = _tiles_from_frame(fr, 96, rng)
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.1,
                                          _lib.default_params(history=history, precond=precond))
    ref = oracle.solve_tiles(d, g, mf, 0.1, f32=True, param=oracle.default_param(m=history))
    f64, gn, un = oracle.eval_cost_f32(u, d, g, mf, 0.1)
    rel = (f64 - ref["cost"]) / numpy.abs(ref["cost"])
    _say("history", history, "max rel", rel.max(), "iters", it.mean(), ref["iters"].mean(),
         "status", numpy.unique(st, return_counts=True))
    assert numpy.all(st == 0)
    assert rel.max() <= 1e-5
    assert numpy.all(gn <= 2e-5 * numpy.maximum(1.0, un))


@pytest.mark.parametrize("precision", ["fp32", "f16split"])
def test_curvature_corrected_h0_saves_iterations(gpu_ctx, precision):
    """The curvature-corrected initial inverse Hessian (precond 2, the
    default) adds the likelihood curvature's variation over the tile to the
    spectral one (precond 1) to first order.  From the plain and from the
    linear start both reach the oracle's minimiser; the corrected one takes
    at least half an iteration fewer on average."""
    from ncs_amd import linstart
    _say("corrected H0,", precision)
    fr = sim.make_frame(11, (256, 256))
    d, g = _tiles_from_frame(fr, 200, numpy.random.default_rng(11))
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    gpu_ctx.set_linear_start(linstart.load())
    ref = oracle.solve_tiles(d, g, mf, 0.1, f32=True)
    prec = _lib.PRECISION_FP32 if precision == "fp32" else _lib.PRECISION_F16_SPLIT
    its = {}
    for start in (_lib.START_PLAIN, _lib.START_LINEAR):
        for pc in (1, 2):
            p = _lib.default_params(history=1, precond=pc, start=start, precision=prec)
            u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.1, p)
            f64, gn, un = oracle.eval_cost_f32(u, d, g, mf, 0.1)
            rel = (f64 - ref["cost"]) / numpy.abs(ref["cost"])
            its[start, pc] = it.mean()
            _say("start", start, "precond", pc, "iterations", it.mean(), "max rel", rel.max())
            assert numpy.all(st == 0)
            assert rel.max() <= 1e-5
            assert numpy.all(gn <= 2e-5 * numpy.maximum(1.0, un))
    for start in (_lib.START_PLAIN, _lib.START_LINEAR):
        assert its[start, 2] < its[start, 1] - 0.5, its


def test_solver_armijo_kernel_semantics(gpu_ctx):
    """The reference kernel's line search and constants (ncs.cl:58-63)."""
    _say("armijo mode")
    fr = sim.make_frame(9, (96, 96))
    d, g = _tiles_from_frame(fr, 64, numpy.random.default_rng(9))
    mf = masks.to_fft_order(masks.gaussian_highpass_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.1, _lib.kernel_semantics_params())
    ref = oracle.solve_tiles(d, g, mf, 0.1, f32=True)
    f64, gn, un = oracle.eval_cost_f32(u, d, g, mf, 0.1)
    rel = (f64 - ref["cost"]) / numpy.abs(ref["cost"])
    _say("armijo rel", rel.max(), "iters", it.mean(), "status", numpy.unique(st, return_counts=True))
    assert numpy.all((st == 0) | (st == -5))
    assert rel.max() <= 1e-4  # epsilon 1e-4 in this mode


def test_fixed_iteration_mode_reaches_at_least_the_oracle_cost(gpu_ctx):
    _say("fixed iteration mode")
    fr = sim.make_frame(10, (96, 96))
    d, g = _tiles_from_frame(fr, 32, numpy.random.default_rng(10))
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    p = _lib.default_params(epsilon=0.0, max_iterations=200)
    u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.1, p)
    ref = oracle.solve_tiles(d, g, mf, 0.1, f32=True)
    f64, _, _ = oracle.eval_cost_f32(u, d, g, mf, 0.1)
    _say("fixed: iters", numpy.unique(it), "status", numpy.unique(st, return_counts=True))
    assert numpy.all(f64 <= ref["cost"] + 1e-5 * numpy.abs(ref["cost"]))
    assert numpy.all((st == -2) | (st == -5) | (st == -4))


def test_status_codes(gpu_ctx):
    _say("status codes")
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    rng = numpy.random.default_rng(12)
    d = rng.uniform(10, 50, (4, 16, 16)).astype(numpy.float32)
    g = rng.uniform(1, 5, (4, 16, 16)).astype(numpy.float32)
    # alpha = 0: the data is the minimiser, converged before any iteration
    u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.0, _lib.default_params())
    assert numpy.all(st == 0) and numpy.all(it == 0) and numpy.array_equal(u, d)
    # one iteration allowed
    u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.1, _lib.default_params(max_iterations=1))
    assert numpy.all(st == _lib.REACHED_MAXITERS) and numpy.all(it == 1)
    # NaN data ends with a failure status instead of looping
    d2 = d.copy()
    d2[1, 3, 3] = numpy.nan
    u, it, st, cost = gpu_ctx.solve_tiles(d2, g, 0.1, _lib.default_params())
    assert st[1] != 0 and st[0] == 0


def test_warm_start_from_the_solution_converges_immediately(gpu_ctx):
    """Device API with a tile start: starting at the solved tiles needs (almost)
    no further iterations, and the fallback keeps the plain start when the
    given start is worse."""
    import torch
    _say("warm start device api")
    fr = sim.make_frame(13, (96, 96))
    d, g = _tiles_from_frame(fr, 64, numpy.random.default_rng(13))
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    p = _lib.default_params()
    u, it, st, cost = gpu_ctx.solve_tiles(d, g, 0.1, p)
    dev = torch.device("cuda:0")
    td, tg = torch.from_numpy(d).to(dev), torch.from_numpy(g).to(dev)
    out = torch.zeros_like(td)
    iters = torch.zeros(64, dtype=torch.int32, device=dev)
    status = torch.zeros_like(iters)
    c2 = torch.zeros(64, dtype=torch.float32, device=dev)
    for start, expect_fast in ((torch.from_numpy(u).to(dev), True),
                               (torch.full_like(td, 1000.0), False)):
        gpu_ctx.solve_tiles_device(td.data_ptr(), tg.data_ptr(), start.data_ptr(), out.data_ptr(),
                                   iters.data_ptr(), status.data_ptr(), c2.data_ptr(), 64, 0.1, p,
                                   torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        it2 = iters.cpu().numpy()
        _say("start", "solution" if expect_fast else "bad", "iters", it2.mean())
        assert numpy.all(status.cpu().numpy() == 0)
        if expect_fast:
            assert it2.mean() <= 2.0
        else:
            assert abs(it2.mean() - it.mean()) <= 1.0  # fell back to the plain start
        rel = numpy.abs(c2.cpu().numpy() - cost) / numpy.abs(cost)
        assert rel.max() <= 1e-5


@pytest.mark.parametrize("factor", [1e-5, 1e-2])
def test_far_warm_start_reaches_the_minimiser(gpu_ctx, factor):
    """A start far below the data with the fallback off: the spectral H0
    models the curvature by its tile mean, which changes by orders of
    magnitude on the way, so a failed line search hands the tile over to
    liblbfgs's scalar H0.  Every tile must converge, to the oracle's cost."""
    import torch
    fr = sim.make_frame(12, (128, 128))
    rng = numpy.random.default_rng(12)
    d, g = _tiles_from_frame(fr, 32, rng)
    d, g = d.astype(numpy.float32), g.astype(numpy.float32)
    mf = masks.to_fft_order(masks.otf_mask()).astype(numpy.float32)
    gpu_ctx.set_otf_mask(mf)
    dev = torch.device("cuda:0")
    td, tg = torch.from_numpy(d).to(dev), torch.from_numpy(g).to(dev)
    tu0 = torch.from_numpy(factor * d).to(dev)
    tu = torch.empty_like(td)
    ti = torch.empty(len(d), dtype=torch.int32, device=dev)
    ts = torch.empty_like(ti)
    p = _lib.default_params(fallback=0, max_iterations=500)
    gpu_ctx.solve_tiles_device(td.data_ptr(), tg.data_ptr(), tu0.data_ptr(), tu.data_ptr(),
                               ti.data_ptr(), ts.data_ptr(), None, len(d), 0.1, p,
                               torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    u, st = tu.cpu().numpy(), ts.cpu().numpy()
    ref = oracle.solve_tiles(d, g, mf, 0.1, f32=True)
    f64, gn, un = oracle.eval_cost_f32(u, d, g, mf, 0.1)
    rel = (f64 - ref["cost"]) / numpy.abs(ref["cost"])
    _say("far start", factor, "status", numpy.unique(st), "iters", ti.float().mean().item(),
         "max rel cost", rel.max())
    assert numpy.all(st == 0)
    assert rel.max() <= 1e-5
