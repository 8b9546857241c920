"""Pins the FP64 conv-net oracle (oracle/convnet_ref.py): torch's CPU conv
(a library routine) for the forward, central finite differences and
torch.autograd in float64 for the backward."""
import numpy
import torch

from oracle import convnet_ref as ref
from ncs_amd import warmstart


def _torch_net(params, x):
    layers = ref.unpack(params)
    a = torch.from_numpy(x)
    out = None
    tl = []
    for li, (w, b) in enumerate(layers):
        tw = torch.tensor(w, requires_grad=True)
        tb = torch.tensor(b, requires_grad=True)
        tl += [tw, tb]
        a = torch.nn.functional.conv2d(a, tw, tb, padding=1)
        if li < 3:
            a = torch.relu(a)
        out = a
    return out[:, 0], tl


def test_forward_matches_torch_conv():
    rng = numpy.random.default_rng(0)
    p = warmstart.init_params(1, last_scale=1.0).astype(numpy.float64)
    x = rng.standard_normal((3, 2, 16, 16))
    y, _ = ref.forward(p, x)
    yt, _ = _torch_net(p, x)
    assert numpy.max(numpy.abs(y - yt.detach().numpy())) < 1e-10 * max(1.0, numpy.max(numpy.abs(y)))


def test_backward_matches_torch_autograd():
    rng = numpy.random.default_rng(1)
    p = warmstart.init_params(2, last_scale=1.0).astype(numpy.float64)
    x = rng.standard_normal((2, 2, 16, 16))
    gy = rng.standard_normal((2, 16, 16))
    y, cache = ref.forward(p, x)
    g = ref.backward(p, cache, gy)
    yt, tl = _torch_net(p, x)
    (yt * torch.from_numpy(gy)).sum().backward()
    gt = numpy.concatenate([t.grad.numpy().ravel() for t in tl])
    assert g.shape == (ref.N_PARAMS,)
    assert numpy.max(numpy.abs(g - gt)) < 1e-9 * max(1.0, numpy.max(numpy.abs(gt)))


def test_backward_finite_differences():
    rng = numpy.random.default_rng(2)
    p = warmstart.init_params(3, last_scale=1.0).astype(numpy.float64)
    x = rng.standard_normal((1, 2, 16, 16))
    gy = rng.standard_normal((1, 16, 16))
    y, cache = ref.forward(p, x)
    g = ref.backward(p, cache, gy)
    h = 1e-6
    for idx in [0, 575, 576, 700, 9830, 12000, 19104, 19392]:
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (numpy.sum(ref.forward(pp, x)[0] * gy) - numpy.sum(ref.forward(pm, x)[0] * gy)) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-5 * max(1.0, abs(fd)), idx


def test_bf16_rounding_is_nearest_even():
    v = numpy.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], numpy.float32)
    r = ref.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


def test_param_layout_matches_header_count():
    assert warmstart.N_PARAMS == ref.N_PARAMS == 19393
    sl = warmstart.layer_slices()
    assert sl[0][0] == slice(0, 576) and sl[-1][1] == slice(19392, 19393)
