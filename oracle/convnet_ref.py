"""FP64 CPU oracle of the warm-start conv net (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward kernels
(ncs_amd/csrc/ncs_convnet.hip).  The net has no counterpart in the
reference; its definition is the one of include/ncs_hip.h: 2 -> 32 -> 32 ->
32 -> 1 channels, 3x3 cross-correlation with zero padding inside the 16 x 16
tile, bias, ReLU after layers 1-3; parameters flat, per layer the weights
[cout][cin][3][3] then the bias.

``bf16=True`` rounds weights, the input and every stored activation to
bfloat16 (round to nearest even) the way the kernels do, keeping FP64
arithmetic otherwise, so kernel-vs-oracle differences come from
accumulation order only.
"""
import numpy

LAYERS = [(2, 32), (32, 32), (32, 32), (32, 1)]
N_PARAMS = 19393


def bf16_round(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = numpy.ascontiguousarray(a, dtype=numpy.float32)
    u = f.view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(numpy.uint32).view(numpy.float32).astype(numpy.float64)


def unpack(params):
    out, off = [], 0
    p = numpy.asarray(params, dtype=numpy.float64)
    for cin, cout in LAYERS:
        nw = cout * cin * 9
        out.append((p[off:off + nw].reshape(cout, cin, 3, 3), p[off + nw:off + nw + cout]))
        off += nw + cout
    return out


def conv3x3(x, w, b):
    """x [n, cin, 16, 16], w [cout, cin, 3, 3] -> [n, cout, 16, 16]."""
    n, cin, H, W = x.shape
    xp = numpy.zeros((n, cin, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    out = numpy.zeros((n, w.shape[0], H, W))
    for ky in range(3):
        for kx in range(3):
            out += numpy.einsum("nchw,oc->nohw", xp[:, :, ky:ky + H, kx:kx + W], w[:, :, ky, kx])
    return out + b[None, :, None, None]


def conv3x3_grads(x, w, gout):
    """Gradients of conv3x3 w.r.t. x, w, b given dL/dout."""
    n, cin, H, W = x.shape
    xp = numpy.zeros((n, cin, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    gw = numpy.zeros_like(w)
    gxp = numpy.zeros_like(xp)
    for ky in range(3):
        for kx in range(3):
            gw[:, :, ky, kx] = numpy.einsum("nohw,nchw->oc", gout, xp[:, :, ky:ky + H, kx:kx + W])
            gxp[:, :, ky:ky + H, kx:kx + W] += numpy.einsum("nohw,oc->nchw", gout, w[:, :, ky, kx])
    return gxp[:, :, 1:-1, 1:-1], gw, gout.sum(axis=(0, 2, 3))


def forward(params, x, bf16=False):
    """y [n, 16, 16] and the cache for backward; x [n, 2, 16, 16]."""
    r = bf16_round if bf16 else (lambda a: numpy.asarray(a, dtype=numpy.float64))
    layers = unpack(params)
    a = r(x)
    cache = [a]
    for li, (w, b) in enumerate(layers):
        z = conv3x3(a, r(w), b)
        if li < len(layers) - 1:
            a = r(numpy.maximum(z, 0.0))
            cache.append(a)
        else:
            y = z[:, 0]
    return y, cache


def backward(params, cache, gy, bf16=False):
    """dL/dparams (flat, FP64) for dL/dy = gy."""
    r = bf16_round if bf16 else (lambda a: numpy.asarray(a, dtype=numpy.float64))
    layers = unpack(params)
    grads = []
    g = numpy.asarray(gy, dtype=numpy.float64)[:, None]
    for li in range(len(layers) - 1, -1, -1):
        w, b = layers[li]
        gx, gw, gb = conv3x3_grads(cache[li], r(w), g)
        grads.append((gw, gb))
        if li > 0:
            g = gx * (cache[li] > 0)
            if bf16:
                g = bf16_round(g)
    out = []
    for gw, gb in reversed(grads):
        out += [gw.ravel(), gb]
    return numpy.concatenate(out)
