"""pyCNCS's Python layer on libncs_hip.so.

Mirrors clib/pyCNCS/ncs_c.py (same names, argument meaning, checks and
exceptions): the NCSCSubRegion class over the per-sub-region C ABI
(ncsSRInitialize ... ncsSRCleanup), checkOTFMask, cReduceNoise (the C
library's whole-image entry, ncsReduceNoise) and pyReduceNoise (Python
tiling, one sub-region solve per window).  All of them run on the GPU:
ncsSRSolve is the FP32 solver with liblbfgs's semantics and the ncsSRCalc*
functions evaluate in FP64 on the device.

Differences from the reference, kept on purpose:
  * cReduceNoise returns a float64 image whatever the input dtype (the
    reference returns zeros for non-float64 input, ncs_c.py:267-276);
  * size errors raise NCSCException (the reference names an undefined
    NCSException there, ncs_c.py:78, 230, 233, 265, which raises NameError);
  * sub-region sizes are even, 4..16.
"""
import numpy

from . import _lib


class NCSCException(Exception):
    pass


def _f64(a):
    return numpy.ascontiguousarray(a, dtype=numpy.float64)


class NCSCSubRegion(object):
    """NCS solver for a square sub-region of an image (ncs_c.py:67-236)."""

    def __init__(self, r_size=None, strict=True, **kwds):
        super().__init__(**kwds)
        self.alpha = None
        self.r_size = r_size
        self.strict = strict
        self.image = None
        if r_size is None or (r_size % 2) != 0:
            raise NCSCException("Sub region size must be divisible by 2!")
        self.lib = _lib.load()
        self.c_ncs = self.lib.ncsSRInitialize(int(r_size))
        if not self.c_ncs:
            raise NCSCException("ncsSRInitialize failed: " + _lib.last_error())

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    def calcCost(self, u):
        """Cost at u (used by pySolve)."""
        self.setU(u.reshape(self.image.shape))
        return self.calcLogLikelihood() + self.alpha * self.calcNoiseContribution()

    def calcCostGradient(self, u):
        """Cost gradient at the u of the last calcCost (used by pySolve)."""
        return self.calcLLGradient() + self.alpha * self.calcNCGradient()

    def calcLLGradient(self):
        gradient = numpy.zeros(self.r_size * self.r_size, dtype=numpy.float64)
        self.lib.ncsSRCalcLLGradient(self.c_ncs, gradient.ctypes.data)
        return gradient

    def calcLogLikelihood(self):
        return self.lib.ncsSRCalcLogLikelihood(self.c_ncs)

    def calcNCGradient(self):
        gradient = numpy.zeros(self.r_size * self.r_size, dtype=numpy.float64)
        self.lib.ncsSRCalcNCGradient(self.c_ncs, gradient.ctypes.data)
        return gradient

    def calcNoiseContribution(self):
        return self.lib.ncsSRCalcNoiseContribution(self.c_ncs)

    def cleanup(self):
        if getattr(self, "c_ncs", None):
            self.lib.ncsSRCleanup(self.c_ncs)
            self.c_ncs = None

    def cSolve(self, alpha, verbose=True):
        """L-BFGS solve from u = image; returns u (ncs_c.py:116-124)."""
        ret = self.lib.ncsSRSolve(self.c_ncs, float(alpha), int(bool(verbose)))
        if verbose:
            print("L-BFGS method returned {0:d}".format(ret))
        if self.strict and (ret != 0):
            raise NCSCException("Solver failed with error code {0:d}!".format(ret))
        return self.getU()

    def getU(self):
        u = numpy.zeros((self.r_size, self.r_size), dtype=numpy.float64)
        self.lib.ncsSRGetU(self.c_ncs, u.ctypes.data)
        return u

    def newRegion(self, image, gamma):
        self.image = image
        if self.strict:
            if image.shape[0] != image.shape[1]:
                raise NCSCException("Sub-region image must be square!")
            if image.size != self.r_size * self.r_size:
                raise NCSCException("Sub-region image size must match sub-region size!")
            if gamma.shape[0] != gamma.shape[1]:
                raise NCSCException("Sub-region gamma must be square!")
            if gamma.size != self.r_size * self.r_size:
                raise NCSCException("Sub-region gamma size must match sub-region size!")
        im, gm = _f64(image), _f64(gamma)
        if im.size != self.r_size * self.r_size or gm.size != im.size:
            raise NCSCException("Sub-region size mismatch!")
        self.lib.ncsSRNewRegion(self.c_ncs, im.ctypes.data, gm.ctypes.data)

    def pySolve(self, alpha, verbose=True):
        """scipy L-BFGS-B without gradients over the GPU cost (the
        reference's test helper, ncs_c.py:152-173)."""
        import scipy.optimize
        self.alpha = alpha
        opts = {"disp": verbose, "maxiter": 200}
        outi = scipy.optimize.minimize(self.calcCost, self.image, method="L-BFGS-B",
                                       options=opts)
        if verbose:
            print("Status", outi.status)
            print("Message", outi.message)
            print("Iterations", outi.nit)
        return outi.x.reshape(self.image.shape)

    def pySolveGradient(self, alpha, verbose=True):
        """scipy L-BFGS-B with the GPU gradient, bounded at 0 (the
        reference's test helper, ncs_c.py:175-206)."""
        import scipy.optimize
        self.alpha = alpha
        opts = {"disp": verbose, "maxiter": 200}
        bounds = scipy.optimize.Bounds(numpy.zeros(self.image.size),
                                       numpy.full(self.image.size, numpy.inf))
        outi = scipy.optimize.minimize(self.calcCost, self.image, bounds=bounds,
                                       jac=self.calcCostGradient, method="L-BFGS-B",
                                       options=opts)
        if verbose:
            print("Status", outi.status)
            print("Message", outi.message)
            print("Iterations", outi.nit)
        return outi.x.reshape(self.image.shape)

    def setOTFMask(self, otf_mask):
        """Centred mask; stored in FFT order (ncs_c.py:208-223)."""
        if self.strict:
            if otf_mask.shape[0] != otf_mask.shape[1]:
                raise NCSCException("OTF must be square!")
            if otf_mask.size != self.r_size * self.r_size:
                raise NCSCException("OTF size must match sub-region size!")
            if not checkOTFMask(otf_mask):
                raise NCSCException("OTF does not have the expected symmetry!")
        tmp = _f64(numpy.fft.fftshift(otf_mask))
        if tmp.size != self.r_size * self.r_size:
            raise NCSCException("OTF size must match sub-region size!")
        self.lib.ncsSRSetOTFMask(self.c_ncs, tmp.ctypes.data)

    def setU(self, u):
        if self.strict:
            if u.shape[0] != u.shape[1]:
                raise NCSCException("u must be square!")
            if u.size != self.r_size * self.r_size:
                raise NCSCException("u size must match sub-region size!")
        uu = _f64(u)
        if uu.size != self.r_size * self.r_size:
            raise NCSCException("u size must match sub-region size!")
        self.lib.ncsSRSetU(self.c_ncs, uu.ctypes.data)


def checkOTFMask(otf_mask):
    """Symmetry test of ncs_c.py:239-247: max(imag(ifft2(fftshift(M)))) <= 1e-12."""
    otf_mask_fft = numpy.fft.ifft2(numpy.fft.fftshift(otf_mask))
    return not (numpy.max(numpy.imag(otf_mask_fft)) > 1.0e-12)


def cReduceNoise(image, gamma, otf_mask, alpha, strict=True):
    """The C library's whole-image entry (ncs_c.py:250-276): FFT-order mask
    of side r (even, 4..16), float64 image out."""
    if strict:
        if otf_mask.shape[0] != otf_mask.shape[1]:
            raise NCSCException("OTF must be square!")
        if (otf_mask.shape[0] % 2) != 0:
            raise NCSCException("Sub region size must be divisible by 2!")
    r = otf_mask.shape[0]
    im = _f64(image)
    if otf_mask.shape != (r, r) or r < 4 or r > 16 or r % 2:
        raise NCSCException("the GPU solver takes even sub-region sizes 4..16")
    if im.ndim != 2 or min(im.shape) < r - 2 or numpy.shape(gamma) != im.shape:
        raise NCSCException("image and gamma must be 2-D of one shape, at least %d pixels "
                            "on each side" % (r - 2))
    # the converted arrays stay bound until the call returns
    gm, mk = _f64(gamma), _f64(otf_mask)
    ncs_image = numpy.zeros(im.shape, dtype=numpy.float64)
    _lib.load().ncsReduceNoise(ncs_image.ctypes.data, im.ctypes.data, gm.ctypes.data,
                               mk.ctypes.data, float(alpha), im.shape[0], im.shape[1], r)
    return ncs_image


def pyReduceNoise(image, gamma, otf_mask, alpha, strict=True):
    """Python tiling with one sub-region solve per window (ncs_c.py:279-322):
    CENTRED mask of side r, windows of side r at stride r - 2 over the image
    padded by one replicated pixel, the last window clamped, later windows
    overwriting overlaps."""
    r_size = otf_mask.shape[0]
    s_size = r_size - 2
    ncs_sr = NCSCSubRegion(r_size=r_size, strict=strict)
    ncs_sr.setOTFMask(otf_mask)
    pad_image = numpy.pad(image, 1, "edge")
    pad_gamma = numpy.pad(gamma, 1, "edge")
    ncs_image = numpy.zeros_like(image)  # the input's dtype, as the reference
    for i in range(0, pad_image.shape[0], s_size):
        bx = pad_image.shape[0] - r_size if (i + r_size) > pad_image.shape[0] else i
        ex = bx + r_size
        for j in range(0, pad_image.shape[1], s_size):
            by = pad_image.shape[1] - r_size if (j + r_size) > pad_image.shape[1] else j
            ey = by + r_size
            ncs_sr.newRegion(pad_image[bx:ex, by:ey], pad_gamma[bx:ex, by:ey])
            im = ncs_sr.cSolve(alpha, verbose=False)
            ncs_image[bx:ex - 2, by:ey - 2] = im[1:-1, 1:-1]
    ncs_sr.cleanup()
    return ncs_image
