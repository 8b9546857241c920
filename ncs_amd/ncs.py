"""reduceNoise drop-in: the host API of pyOpenCLNCS / pyCUDANCS on gfx950.

Mirrors opencl/pyOpenCLNCS/ncs.py (and cuda/pyCUDANCS/ncs.py, identical
API): ``reduceNoise(images, gamma, otf_mask, alpha, strict=True,
verbose=False)`` takes a list of photo-electron images, the read-noise map
gamma = variance / gain^2, a CENTRED 16 x 16 OTF mask and alpha, and returns
the list of corrected float32 images.  Tiling (16 x 16 windows, stride 14,
1-px replicated halo, last window clamped, later windows owning overlaps),
the L-BFGS solve of every window and the stitch all run in one persistent
HIP launch (libncs_hip.so); Python only marshals arrays.

Extensions (keyword-only, defaults reproduce the reference):
  * gamma may be a list (one map per frame) and frames may differ in size;
  * alpha may be a list (one value per frame);
  * params: solver parameters (see ncs_amd._lib.default_params);
  * geometry: "reference" (default) or "blocks" (non-overlapping 16 x 16
    blocks, zero-padded edges excluded from the likelihood);
  * return_info=True also returns per-tile iterations / status / final cost.
"""
import warnings

import numpy

from . import _lib

# failure names of opencl/pyOpenCLNCS/ncs.py:170-175
FAILURE_TYPES = {-1: "Unstarted",
                 -2: "Reached maximum iterations",
                 -3: "Increasing gradient",
                 -4: "Reached minimum step size",
                 -5: "Reached machine precision"}


class NCSHipException(Exception):
    pass


# names used by the reference hosts
NCSOpenCLException = NCSHipException
NCSCUDAException = NCSHipException


def checkOTFMask(otf_mask):
    """Symmetry test of opencl/pyOpenCLNCS/ncs.py:225-233 (and
    clib/pyCNCS/ncs_c.py:239-247): max(imag(ifft2(fftshift(M)))) <= 1e-12."""
    otf_mask_fft = numpy.fft.ifft2(numpy.fft.fftshift(otf_mask))
    return not (numpy.max(numpy.imag(otf_mask_fft)) > 1.0e-12)


def _geometry_code(geometry):
    if geometry in ("reference", _lib.GEOM_REFERENCE):
        return _lib.GEOM_REFERENCE
    if geometry in ("blocks", _lib.GEOM_BLOCKS):
        return _lib.GEOM_BLOCKS
    raise ValueError("geometry must be 'reference' or 'blocks'")


class NCSHip(object):
    """The class of opencl/pyOpenCLNCS/ncs.py:39-222 on libncs_hip."""

    def __init__(self, strict=True, device=0, params=None, size=16, **kwds):
        super().__init__(**kwds)
        if size < 4 or size > 16 or size % 2:
            raise NCSHipException("the GPU solver takes even sub-region sizes 4..16")
        self.size = size
        self.strict = strict
        self.ctx = _lib.context(device)
        self.params = params if params is not None else _lib.default_params()
        self.gamma = None
        self.otf_mask = None
        self.last_info = None

    def setGamma(self, gamma):
        """One map for all images (ncs.py:200-206) or a list, one per image."""
        if isinstance(gamma, (list, tuple)):
            self.gamma = [numpy.asarray(g, dtype=numpy.float32) for g in gamma]
        else:
            self.gamma = numpy.asarray(gamma).astype(numpy.float32)

    def setOTFMask(self, otf_mask):
        """Centred mask; checks as ncs.py:208-222, stored in FFT order."""
        otf_mask = numpy.asarray(otf_mask)
        if self.strict:
            if otf_mask.ndim != 2 or otf_mask.shape[0] != otf_mask.shape[1]:
                raise NCSHipException("OTF must be square!")
            if otf_mask.size != self.size * self.size:
                raise NCSHipException("OTF size must match sub-region size!")
            if not checkOTFMask(otf_mask):
                raise NCSHipException("OTF does not have the expected symmetry!")
        elif otf_mask.shape != (self.size, self.size):
            raise NCSHipException("OTF size must match sub-region size!")
        self.otf_mask = numpy.fft.fftshift(otf_mask).astype(numpy.float32)

    def reduceNoise(self, images, alpha, verbose=False, geometry="reference",
                    return_info=False):
        """Run NCS on a list of images (ncs.py:47-198)."""
        if self.otf_mask is None or self.gamma is None:
            raise NCSHipException("set the OTF mask and gamma first")
        if isinstance(images, numpy.ndarray) and images.ndim == 2:
            images = [images]
        images = list(images)
        n = len(images)
        if n == 0:
            return ([], {}) if return_info else []
        geom = _geometry_code(geometry)
        gammas = self.gamma if isinstance(self.gamma, list) else [self.gamma] * n
        if len(gammas) != n:
            raise NCSHipException("one gamma map per image is required")
        alphas = list(alpha) if hasattr(alpha, "__len__") else [alpha] * n
        if len(alphas) != n:
            raise NCSHipException("one alpha per image is required")
        datas, outs = [], []
        for h in range(n):
            im = numpy.asarray(images[h])
            if im.ndim != 2:
                raise NCSHipException("images must be 2-D")
            if gammas[h].shape != im.shape:
                raise NCSHipException("gamma must have the shape of image %d" % h)
            if geom == _lib.GEOM_REFERENCE and min(im.shape) < self.size - 2:
                raise NCSHipException("images must be at least %d pixels on each side"
                                      % (self.size - 2))
            datas.append(im.astype(numpy.float32))
            outs.append(numpy.zeros(im.shape, dtype=numpy.float32))
        self.ctx.set_otf_mask(self.otf_mask)
        iters, status, cost = self.ctx.reduce_frames(datas, gammas, outs, alphas, geom,
                                                     self.params)
        # status handling of ncs.py:156-184
        failures = {}
        n_fails = int(numpy.count_nonzero(status != 0))
        if n_fails > 0:
            n_maxp = int(numpy.count_nonzero(status == -5))
            if n_maxp != n_fails:
                warnings.warn("Noise reduction failed on {0:d} sub-regions.".format(n_fails))
            for i in range(-5, 0):
                nf = int(numpy.count_nonzero(status == i))
                if nf != 0:
                    failures[i] = nf
        if verbose:
            print("Processed {0:d} sub-regions.".format(len(status)))
            print("Minimum iterations: {0:d}".format(int(numpy.min(iters))))
            print("Maximum iterations: {0:d}".format(int(numpy.max(iters))))
            print("Median iterations: {0:.3f}".format(float(numpy.median(iters))))
            for key in failures:
                print(failures[key], "failures of type '" + FAILURE_TYPES[key] + "'")
            print()
        info = dict(iterations=iters, status=status, cost=cost, failures=failures)
        self.last_info = info
        if return_info:
            return outs, info
        return outs


def reduceNoise(images, gamma, otf_mask, alpha, strict=True, verbose=False, *,
                params=None, geometry="reference", return_info=False, device=0):
    """Run NCS on images with libncs_hip (opencl/pyOpenCLNCS/ncs.py:236-251).

    images: list of 2-D arrays in photo-electrons; gamma: one map or one per
    image; otf_mask: 16 x 16 CENTRED; alpha: scalar or one per image.
    Returns a list of float32 images (and the per-tile info dict when
    return_info=True)."""
    ncs = NCSHip(strict=strict, device=device, params=params)
    ncs.setOTFMask(otf_mask)
    ncs.setGamma(gamma)
    return ncs.reduceNoise(images, alpha, verbose=verbose, geometry=geometry,
                           return_info=return_info)


def cReduceNoise(image, gamma, otf_mask, alpha, strict=True, *, params=None, device=0):
    """Twin of clib/pyCNCS/ncs_c.py:250-276 on the GPU: FFT-order mask in,
    float64 image out (FP32 solve)."""
    otf_mask = numpy.asarray(otf_mask)
    if strict:
        if otf_mask.shape[0] != otf_mask.shape[1]:
            raise NCSHipException("OTF must be square!")
        if (otf_mask.shape[0] % 2) != 0:
            raise NCSHipException("Sub region size must be divisible by 2!")
    r = otf_mask.shape[0]
    if otf_mask.shape != (r, r) or r < 4 or r > 16 or r % 2:
        raise NCSHipException("the GPU solver takes even sub-region sizes 4..16")
    ncs = NCSHip(strict=False, device=device, params=params, size=r)
    ncs.setGamma(numpy.asarray(gamma))
    ncs.setOTFMask(numpy.fft.ifftshift(otf_mask))
    out = ncs.reduceNoise([numpy.asarray(image)], alpha)
    return out[0].astype(numpy.float64)


def solveTiles(data, gamma, otf_mask_fft, alpha, params=None, device=0):
    """The reference kernel's own batch interface (opencl/ncs.cl:820-826):
    independent tiles [n, 16, 16] and an FFT-order mask; returns
    (u, iterations, status, final_cost)."""
    ctx = _lib.context(device)
    ctx.set_otf_mask(numpy.asarray(otf_mask_fft, dtype=numpy.float32).reshape(16, 16))
    return ctx.solve_tiles(data, gamma, alpha,
                           params if params is not None else _lib.default_params())
