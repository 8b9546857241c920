"""CPU checks of the oracle's camera restatement and of pyNCS's tiling.

genmaps (matlab/genmaps.m:29-62) is pinned by values read off the MATLAB
masks by hand: ties at the division points, division points that descend,
and the maps a pixel keeps from the previous frame when no level covers it.
The example scene of matlab/multimaps_example.m is checked by radius.
segpadimg/stitchpadimg (python3-6/pyNCS/denoisetools.py:27-86) are pinned
against the reference C tiling (clib/pyCNCS/ncs.c:71-113, window starts
-1, 13, 27, ... with the edge clamped), gathered here directly."""
import numpy

from oracle import camera


def _maps(shape, values):
    return numpy.stack([numpy.full(shape, v, numpy.float64) for v in values])


def test_genmaps_ties_at_the_division_points():
    ims = numpy.array([[[1499, 1500, 2999, 3000, 65535, 0]]], numpy.float64)
    g, v = camera.genmaps(ims, _maps((1, 6), [1.5, 2, 3]), _maps((1, 6), [10, 100, 1000]),
                          [1500, 3000])
    # ADU < 1500 -> map 1; 1500 <= ADU < 3000 -> map 2; ADU >= 3000 -> map 3
    assert g[0, 0].tolist() == [1.5, 2, 2, 3, 3, 1.5]
    assert v[0, 0].tolist() == [10, 100, 100, 1000, 1000, 10]


def test_genmaps_descending_points_follow_the_mask_order():
    # map 1 where ADU < 3000, then map 3 where ADU >= 1500 (the middle mask,
    # ADU >= 3000 and < 1500, is empty)
    ims = numpy.array([[[0, 1499, 1500, 2999, 3000, 5000]]], numpy.float64)
    g, _ = camera.genmaps(ims, _maps((1, 6), [1.5, 2, 3]), _maps((1, 6), [1, 2, 3]),
                          [3000, 1500])
    assert g[0, 0].tolist() == [1.5, 1.5, 3, 3, 3, 3]


def test_genmaps_uncovered_pixels_keep_the_previous_frame():
    # three maps but three division points: ADU >= 4000 is in no mask
    ims = numpy.array([[[5000.0]], [[100.0]], [[5000.0]], [[3500.0]]])
    g, v = camera.genmaps(ims, _maps((1, 1), [1.5, 2, 3]), _maps((1, 1), [10, 100, 1000]),
                          [1500, 3000, 4000])
    assert g[:, 0, 0].tolist() == [0.0, 1.5, 1.5, 3.0]
    assert v[:, 0, 0].tolist() == [0.0, 10.0, 10.0, 1000.0]


def test_genmaps_multimaps_example_scene_by_radius():
    R = 128
    x = numpy.arange(-R // 2 + 1, R // 2 + 1)
    X, Y = numpy.meshgrid(x, x)
    rr = X * X + Y * Y  # integer squared radius
    Z = numpy.sqrt(rr) / R * numpy.sqrt(2)
    scales = [3500, 4000, 4500, 5000]
    ims = numpy.stack([Z * s for s in scales])
    g, v = camera.genmaps(ims, _maps((R, R), [1.5, 2, 3]), _maps((R, R), [10, 100, 1000]),
                          [1500, 3000])
    for k, s in enumerate(scales):
        # ADU = s sqrt(2 rr)/R >= p  <=>  2 s^2 rr >= (p R)^2, in integers
        lvl = (2 * s * s * rr >= (1500 * R) ** 2).astype(int) + \
              (2 * s * s * rr >= (3000 * R) ** 2).astype(int)
        assert numpy.array_equal(g[k], numpy.array([1.5, 2, 3])[lvl]), s
        assert numpy.array_equal(v[k], numpy.array([10, 100, 1000])[lvl]), s
    # every level occurs in the brightest frame
    assert set(numpy.unique(g[3]).tolist()) == {1.5, 2.0, 3.0}


def test_adu_to_photons_clamps_and_gamma():
    d = camera.adu_to_photons(numpy.array([90, 100, 104], numpy.uint16), 100.0, 2.0)
    assert d.dtype == numpy.float32 and d.tolist() == [numpy.float32(1e-6), numpy.float32(1e-6), 2.0]
    assert camera.gamma_from_camera(numpy.float32(8.0), numpy.float32(2.0)) == 2.0


def _reference_windows(img, r=16):
    """Windows of the reference tiling: starts -1, 13, ..., the last one
    clamped to H + 1 - r and the loop ending after the window at that
    position; pixels outside the image replaced by the edge
    (ncs.c:71-113, 136-142)."""
    H = img.shape[0]
    starts = []
    i = -1
    while True:
        s = min(i, H + 1 - r)
        starts.append(s)
        if s != i or i + r >= H + 1:
            break
        i += r - 2
    idx = lambda s: numpy.clip(numpy.arange(s, s + r), 0, H - 1)
    return starts, {(a, b): img[numpy.ix_(idx(a), idx(b))] for a in starts for b in starts}


def test_segpadimg_r14_windows_are_the_reference_windows():
    img = numpy.random.default_rng(3).uniform(0, 100, (112, 112))
    segs = camera.segpadimg(img, 14)
    assert segs.shape == (64, 16, 16)
    starts, ref = _reference_windows(img)
    # eight windows per axis in both: the eighth ends exactly one pixel past
    # the edge, and the reference stops after the window at that position
    assert starts == [-1 + 14 * k for k in range(8)]
    for bj in range(8):
        for bi in range(8):
            assert numpy.array_equal(segs[bj * 8 + bi], ref[(starts[bi], starts[bj])]), (bi, bj)


def test_segpadimg_blocks_are_reference_windows_of_side_r1_plus_2():
    """pyNCS's blocks of R1 = 8 and 10 (windows of 10 and 12) are the
    reference tiling's windows of side R1 + 2 whenever R1 divides the side:
    the frame path with that window side serves them."""
    for r1, R in ((8, 48), (10, 60), (10, 30)):
        img = numpy.random.default_rng(r1 + R).uniform(0, 100, (R, R))
        segs = camera.segpadimg(img, r1)
        ns = R // r1
        starts, ref = _reference_windows(img, r1 + 2)
        assert starts == [-1 + r1 * k for k in range(ns)], (r1, R)
        for bj in range(ns):
            for bi in range(ns):
                assert numpy.array_equal(segs[bj * ns + bi], ref[(starts[bi], starts[bj])])


def test_stitchpadimg_places_interiors():
    img = numpy.random.default_rng(4).uniform(0, 100, (60, 60))
    assert numpy.array_equal(camera.stitchpadimg(camera.segpadimg(img, 10)), img)
    segs = camera.segpadimg(img, 10)
    segs[7, 1:-1, 1:-1] = -1.0  # window 7: column block 1, row block 1
    out = camera.stitchpadimg(segs)
    assert numpy.all(out[10:20, 10:20] == -1.0) and numpy.sum(out == -1.0) == 100
