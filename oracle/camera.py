"""CPU restatement of the camera side of the path and of pyNCS's tiling.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker).

- ``adu_to_photons``: python3-6/pyNCS/denoisetools.py:218-224 (and
  python3-6/NCSdemo_experiment.py:42-49), in float32 as the GPU computes it;
  ``gamma_from_camera``: denoisetools.py:266.
- ``genmaps``: matlab/genmaps.m:29-62, per-pixel gain and variance maps of a
  camera with several gain amplifiers, chosen by each frame's ADU level.
- ``segpadimg`` / ``stitchpadimg``: python3-6/pyNCS/denoisetools.py:27-53 and
  76-86 (matlab/segpadimg.m, stitchpadimg.m), pyNCS's split of a square
  image into R1 x R1 blocks with a one-pixel halo of neighbour pixels.
"""
import numpy


def adu_to_photons(adu, offset, gain):
    """(ADU - offset) / gain in float32, values <= 0 set to 1e-6
    (denoisetools.py:222-224)."""
    d = (numpy.asarray(adu).astype(numpy.float32) - numpy.float32(offset)) / numpy.float32(gain)
    d = numpy.asarray(d, dtype=numpy.float32)
    d[d <= 0] = numpy.float32(1e-6)
    return d


def gamma_from_camera(var, gain):
    """variance / gain^2 in float32 (denoisetools.py:266)."""
    g = numpy.asarray(gain, dtype=numpy.float32)
    return (numpy.asarray(var, dtype=numpy.float32) / (g * g)).astype(numpy.float32)


def genmaps(ims, gain, var, lightlevel):
    """Gain and variance stacks [N][R][C] for raw frames ims [N][R][C], from
    gain/var stacks [NL][R][C] and the NL - 1 division points lightlevel,
    as genmaps.m:29-62 builds them (frames first here, MATLAB keeps them
    third).  The MATLAB loop is kept: level 1 where ADU < lightlevel(1),
    level jj+1 where lightlevel(jj) <= ADU < lightlevel(jj+1), and the last
    level where ADU >= its division point, each mask overwriting the earlier
    ones; maps not covered by any mask keep the previous frame's values
    (zeros before the first frame), since gain0/var0 persist across frames
    (genmaps.m:33-34)."""
    ims = numpy.asarray(ims, dtype=numpy.float64)
    gain = numpy.asarray(gain)
    var = numpy.asarray(var)
    ll = numpy.atleast_1d(numpy.asarray(lightlevel, dtype=numpy.float64))
    n, nl = ims.shape[0], gain.shape[0]
    gain0 = numpy.zeros(ims.shape[1:], dtype=gain.dtype)
    var0 = numpy.zeros(ims.shape[1:], dtype=var.dtype)
    gs = numpy.zeros(ims.shape, dtype=gain.dtype)
    vs = numpy.zeros(ims.shape, dtype=var.dtype)
    for ii in range(n):
        im = ims[ii]
        m = im < ll[0]
        gain0[m] = gain[0][m]
        var0[m] = var[0][m]
        for jj in range(1, nl):  # MATLAB jj = 1 .. NL-1, maps jj+1
            if ll.size >= jj + 1:
                m = (im >= ll[jj - 1]) & (im < ll[jj])
            else:
                m = im >= ll[jj - 1]
            gain0[m] = gain[jj][m]
            var0[m] = var[jj][m]
        gs[ii] = gain0
        vs[ii] = var0
    return gs, vs


def segpadimg(img, r1):
    """(Ns*Ns, r1+2, r1+2) windows of a square image of side R (Ns = R//r1):
    block (bi, bj) covers rows/columns [bi*r1, bi*r1 + r1) plus one pixel of
    its neighbours on each side, the image edge replicated outside
    (denoisetools.py:27-43).  Windows are ordered column block first:
    index bj*Ns + bi (segimg, denoisetools.py:45-53)."""
    img = numpy.asarray(img)
    R = img.shape[0]
    ns = R // r1
    # the case pyNCS tiles: a square image of two or more whole blocks
    assert img.shape == (R, R) and R % r1 == 0 and ns >= 2, "square, r1 | R, R >= 2 r1"

    def strips(a, axis):
        parts = [numpy.take(a, numpy.arange(0, r1 + 1), axis=axis)]
        for ii in range(1, ns - 1):
            parts.append(numpy.take(a, numpy.arange(ii * r1 - 1, (ii + 1) * r1 + 1), axis=axis))
        parts.append(numpy.take(a, numpy.arange((ns - 1) * r1 - 1, a.shape[axis]), axis=axis))
        return numpy.concatenate(parts, axis=axis)

    ims2 = numpy.pad(strips(strips(img, 1), 0), ((1, 1), (1, 1)), mode="edge")
    w = r1 + 2
    n2 = ims2.shape[0] // w
    out = numpy.zeros((n2 * n2, w, w), dtype=numpy.float64)
    for ii in range(n2):
        out[ii * n2:(ii + 1) * n2] = ims2[:, ii * w:(ii + 1) * w].reshape(n2, w, w)
    return out


def stitchpadimg(segs):
    """Inverse of segpadimg on the interiors: window k's [1:-1, 1:-1] goes to
    row block k % Ns, column block k // Ns (denoisetools.py:76-86)."""
    segs = numpy.asarray(segs)
    n = segs.shape[0]
    ns = int(numpy.sqrt(n))
    cols = []
    for cb in range(ns):
        cols.append(numpy.vstack([segs[cb * ns + rb, 1:-1, 1:-1] for rb in range(ns)]))
    return numpy.hstack(cols)
