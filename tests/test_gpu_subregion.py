"""GPU checks of the per-sub-region C ABI (ncsSR*) and pyCNCS's Python layer.

The argtypes table of clib/pyCNCS/ncs_c.py:18-60 is restated here and
applied to libncs_hip.so, the way pyCNCS's loader would bind it
(loadclib.py:14-63 loads the library by name, ncs_c.py sets the types).
Through that binding:
  * the ncsSRCalc* values and gradients (FP64 on the GPU) against the FP64
    oracle's, which numpy.fft pins (tests/test_oracle.py), to 1e-11;
  * ncsSRSolve against the oracle's liblbfgs restatement: return codes, and
    the final FP64 cost within 1e-5 relative;
  * cReduceNoise and pyReduceNoise of ncs_amd.ncs_c against the oracle's
    whole-image path on the reference's own test inputs
    (clib/pyCNCS/test/test_im.py:12-38: image U(0.01, 10), gamma U(2, 4),
    alpha 0.02, a random 10 x 10 mask, 144 crops).
"""
import ctypes

import numpy
import pytest
from numpy.ctypeslib import ndpointer

import oracle
from ncs_amd import _lib, masks, sim
from ncs_amd import ncs_c

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _bind_like_pycncs():
    """libncs_hip.so with the argtypes / restypes of ncs_c.py:18-60."""
    _lib.load()  # torch's HIP runtime first, as the package does
    ncs = ctypes.cdll.LoadLibrary(_lib.LIB_PATH)
    f64 = ndpointer(dtype=numpy.float64)
    ncs.ncsReduceNoise.argtypes = [f64, f64, f64, f64, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int]
    ncs.ncsSRCalcLLGradient.argtypes = [ctypes.c_void_p, f64]
    ncs.ncsSRCalcLogLikelihood.argtypes = [ctypes.c_void_p]
    ncs.ncsSRCalcLogLikelihood.restype = ctypes.c_double
    ncs.ncsSRCalcNCGradient.argtypes = [ctypes.c_void_p, f64]
    ncs.ncsSRCalcNoiseContribution.argtypes = [ctypes.c_void_p]
    ncs.ncsSRCalcNoiseContribution.restype = ctypes.c_double
    ncs.ncsSRCleanup.argtypes = [ctypes.c_void_p]
    ncs.ncsSRGetU.argtypes = [ctypes.c_void_p, f64]
    ncs.ncsSRInitialize.argtypes = [ctypes.c_int]
    ncs.ncsSRInitialize.restype = ctypes.c_void_p
    ncs.ncsSRNewRegion.argtypes = [ctypes.c_void_p, f64, f64]
    ncs.ncsSRSetOTFMask.argtypes = [ctypes.c_void_p, f64]
    ncs.ncsSRSetU.argtypes = [ctypes.c_void_p, f64]
    ncs.ncsSRSolve.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    ncs.ncsSRSolve.restype = ctypes.c_int
    return ncs


@pytest.fixture(scope="module")
def pycncs(gpu_ctx):
    return _bind_like_pycncs()


def _c(a):
    return numpy.ascontiguousarray(a, dtype=numpy.float64)


@pytest.mark.parametrize("r", [16, 10])
@pytest.mark.parametrize("kind", ["otf", "random", "asymmetric"])
def test_calc_functions_match_the_fp64_oracle(pycncs, r, kind):
    """LL, its gradient, NC and its gradient at random u (the inputs of
    clib/pyCNCS/test/test_sr.py:12-96).  The asymmetric mask fails
    checkOTFMask (strict=False only); there both sides apply FFTW's c2r
    conventions to the half spectrum, a case the reference's own tests
    leave open, so it is a consistency check between the two restatements."""
    _say("calc", r, kind)
    rng = numpy.random.default_rng(100 + r)
    if kind == "otf":
        mask_fft = masks.to_fft_order(masks.otf_mask(r=r))
    elif kind == "random":
        mask_fft = masks.to_fft_order(masks.random_mask(r, rng))
    else:
        mask_fft = rng.uniform(0.0, 1.0, (r, r))
    sr = pycncs.ncsSRInitialize(r)
    ref = oracle.NCSCSubRegion(r_size=r, strict=False)
    assert sr
    try:
        ref.setOTFMask(numpy.fft.ifftshift(mask_fft))  # takes it centred
        pycncs.ncsSRSetOTFMask(sr, _c(mask_fft))
        worst = 0.0
        for _ in range(5):
            gamma = rng.uniform(2.0, 4.0, (r, r))
            image = rng.uniform(0.01, 10.0, (r, r))
            u = rng.uniform(0.01, 10.0, (r, r))
            u[0, 0] = -5.0  # the negative-value guard
            pycncs.ncsSRNewRegion(sr, _c(image), _c(gamma))
            pycncs.ncsSRSetU(sr, _c(u))
            ref.newRegion(image, gamma)
            ref.setU(u)
            ll, llo = pycncs.ncsSRCalcLogLikelihood(sr), ref.calcLogLikelihood()
            nc, nco = pycncs.ncsSRCalcNoiseContribution(sr), ref.calcNoiseContribution()
            g1, g2 = numpy.zeros(r * r), numpy.zeros(r * r)
            pycncs.ncsSRCalcLLGradient(sr, g1)
            pycncs.ncsSRCalcNCGradient(sr, g2)
            g1o, g2o = ref.calcLLGradient(), ref.calcNCGradient()
            errs = [abs(ll - llo) / abs(llo), abs(nc - nco) / abs(nco),
                    numpy.max(numpy.abs(g1 - g1o)) / numpy.max(numpy.abs(g1o)),
                    numpy.max(numpy.abs(g2 - g2o)) / numpy.max(numpy.abs(g2o))]
            worst = max(worst, max(errs))
            u_back = numpy.zeros((r, r))
            pycncs.ncsSRGetU(sr, u_back)
            assert numpy.array_equal(u_back, u)
        _say("  worst relative error", worst)
        assert worst <= 1e-11
    finally:
        pycncs.ncsSRCleanup(sr)
        ref.cleanup()


@pytest.mark.parametrize("r", [16, 10])
def test_subregion_solve_matches_the_oracle(pycncs, r):
    """ncsSRSolve on windows of a synthetic filament frame (OTF mask) and on
    the reference's random-mask inputs: return code 0 as the oracle's, final
    FP64 cost within 1e-5 relative of the oracle's minimiser."""
    _say("solve", r)
    fr = sim.make_frame(40 + r, (64, 64))
    rng = numpy.random.default_rng(r)
    cases = []
    mask_otf = masks.to_fft_order(masks.otf_mask(r=r))
    for k in range(6):
        y, x = rng.integers(0, 64 - r, 2)
        cases.append((fr["data"][y:y + r, x:x + r], fr["gamma"][y:y + r, x:x + r],
                      mask_otf, 0.1))
    mask_rand = masks.to_fft_order(masks.random_mask(r, rng))
    for k in range(6):
        cases.append((rng.uniform(0.01, 10.0, (r, r)), rng.uniform(2.0, 4.0, (r, r)),
                      mask_rand, 0.02))
    sr = pycncs.ncsSRInitialize(r)
    osr = oracle.NCSCSubRegion(r_size=r, strict=False)
    try:
        worst = 0.0
        for image, gamma, mask_fft, alpha in cases:
            pycncs.ncsSRSetOTFMask(sr, _c(mask_fft))
            pycncs.ncsSRNewRegion(sr, _c(image), _c(gamma))
            ret = pycncs.ncsSRSolve(sr, alpha, 0)
            u = numpy.zeros((r, r))
            pycncs.ncsSRGetU(sr, u)
            osr.setOTFMask(numpy.fft.ifftshift(mask_fft))
            osr.newRegion(_c(image), _c(gamma))
            uo = osr.cSolve(alpha)
            ro = osr.last_status
            f, _ = oracle.cost_grad(u, image, gamma, mask_fft, alpha)
            fo, _ = oracle.cost_grad(uo, image, gamma, mask_fft, alpha)
            worst = max(worst, (f - fo) / abs(fo))
            assert ret == 0 and ro == 0, (ret, ro)
        _say("  worst cost rel", worst)
        assert worst <= 1e-5
    finally:
        pycncs.ncsSRCleanup(sr)
        osr.cleanup()


def test_ncscsubregion_class_on_the_gpu(gpu_ctx):
    """ncs_amd.ncs_c.NCSCSubRegion (centred mask, checks, exceptions) against
    the oracle's class: calc values, cSolve, and the strict error paths."""
    r = 16
    rng = numpy.random.default_rng(5)
    mask = masks.otf_mask(r=r)
    image = rng.uniform(0.01, 10.0, (r, r))
    gamma = rng.uniform(2.0, 4.0, (r, r))
    a = ncs_c.NCSCSubRegion(r_size=r)
    b = oracle.NCSCSubRegion(r_size=r)
    try:
        for s in (a, b):
            s.setOTFMask(mask)
            s.newRegion(image, gamma)
            s.alpha = 0.1
        u = rng.uniform(0.01, 10.0, (r, r))
        assert numpy.isclose(a.calcCost(u), _oracle_cost(b, u), rtol=1e-11)
        ua = a.cSolve(0.1, verbose=False)
        ub = b.cSolve(0.1)
        fa, _ = oracle.cost_grad(ua, image, gamma, masks.to_fft_order(mask), 0.1)
        fb, _ = oracle.cost_grad(ub, image, gamma, masks.to_fft_order(mask), 0.1)
        _say("class: cost rel", (fa - fb) / abs(fb))
        assert (fa - fb) / abs(fb) <= 1e-5
        with pytest.raises(ncs_c.NCSCException):
            a.setOTFMask(numpy.ones((r, r + 1)))
        with pytest.raises(ncs_c.NCSCException):
            a.newRegion(numpy.ones((8, 8)), numpy.ones((8, 8)))
        with pytest.raises(ncs_c.NCSCException):
            a.setOTFMask(numpy.random.default_rng(1).uniform(0, 1, (r, r)))  # not symmetric
    finally:
        a.cleanup()
        b.cleanup()


def _oracle_cost(sr, u):
    sr.setU(u)
    return sr.calcLogLikelihood() + sr.alpha * sr.calcNoiseContribution()


def test_pyreducenoise_and_creducenoise_on_the_reference_test_inputs(gpu_ctx):
    """clib/pyCNCS/test/test_im.py:12-38 on the GPU: pyReduceNoise (Python
    tiling, ncsSRSolve per window) and cReduceNoise (ncsReduceNoise) on 144
    crops of a 30 x 30 image with a random 10 x 10 mask, each against the
    oracle's cReduceNoise (FP64) at the reference GPU test's bar, atol 0.02
    (opencl/pyOpenCLNCS/test/test_im.py:38-44)."""
    im_size, r, alpha = 30, 10, 0.02
    rng = numpy.random.default_rng(12)
    gamma = rng.uniform(2.0, 4.0, (im_size, im_size))
    image = rng.uniform(0.01, 10.0, (im_size, im_size))
    otfmask = masks.random_mask(r, rng)
    shift = numpy.fft.fftshift(otfmask)
    worst_py = worst_c = 0.0
    for ix in range(0, 12, 3):
        for iy in range(0, 12, 3):
            im, gm = image[ix:, iy:], gamma[ix:, iy:]
            ref = oracle.cReduceNoise(im, gm, shift, alpha)
            py = ncs_c.pyReduceNoise(im, gm, otfmask, alpha)
            cc = ncs_c.cReduceNoise(im, gm, shift, alpha)
            assert py.shape == cc.shape == ref.shape
            worst_py = max(worst_py, float(numpy.max(numpy.abs(py - ref))))
            worst_c = max(worst_c, float(numpy.max(numpy.abs(cc - ref))))
    _say("pyReduceNoise max |d| %.2e, cReduceNoise max |d| %.2e" % (worst_py, worst_c))
    assert worst_py <= 0.02 and worst_c <= 0.02
