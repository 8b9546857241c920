"""FP64 CPU oracle for the NCS noise-correction path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg, always as the checker.  The product
package ``ncs_amd`` never imports it.

The C library ``libncs_oracle.so`` (built from ncs_oracle.c by the Makefile
next to this file) restates clib/pyCNCS/ncs.c in double precision; its
header ncs_oracle.h names the two third-party libraries the reference links
(FFTW3, liblbfgs 1.10) and how each is restated.

This module also mirrors the pyCNCS Python API (clib/pyCNCS/ncs_c.py):
``NCSCSubRegion``, ``checkOTFMask``, ``cReduceNoise`` and ``pyReduceNoise``.
"""
import ctypes
import os

import numpy
from numpy.ctypeslib import ndpointer

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libncs_oracle.so")

# liblbfgs codes used below (clib/pyCNCS/lbfgs.h:72-147)
LBFGS_SUCCESS = 0
LBFGS_ALREADY_MINIMIZED = 2
LBFGSERR_ROUNDING_ERROR = -1001
LBFGSERR_MINIMUMSTEP = -1000
LBFGSERR_MAXIMUMLINESEARCH = -998
LBFGSERR_MAXIMUMITERATION = -997
LBFGSERR_WIDTHTOOSMALL = -996
LBFGSERR_INCREASEGRADIENT = -994


class NCSOParam(ctypes.Structure):
    """Mirror of ``ncso_param`` (subset of lbfgs_parameter_t, lbfgs.h:198-358)."""
    _fields_ = [("m", ctypes.c_int),
                ("epsilon", ctypes.c_double),
                ("past", ctypes.c_int),
                ("delta", ctypes.c_double),
                ("max_iterations", ctypes.c_int),
                ("max_linesearch", ctypes.c_int),
                ("min_step", ctypes.c_double),
                ("max_step", ctypes.c_double),
                ("ftol", ctypes.c_double),
                ("gtol", ctypes.c_double),
                ("xtol", ctypes.c_double)]


_lib = None


def load():
    """Load (building once with make if needed) the oracle library."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", _HERE])
    lib = ctypes.cdll.LoadLibrary(_LIB_PATH)
    f64 = ndpointer(dtype=numpy.float64, flags="C_CONTIGUOUS")
    f32 = ndpointer(dtype=numpy.float32, flags="C_CONTIGUOUS")
    vp = ctypes.c_void_p
    lib.ncso_param_init.argtypes = [ctypes.POINTER(NCSOParam)]
    lib.ncsReduceNoise.argtypes = [f64, f64, f64, f64, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ncsSRInitialize.argtypes = [ctypes.c_int]
    lib.ncsSRInitialize.restype = vp
    lib.ncsSRCleanup.argtypes = [vp]
    lib.ncsSRNewRegion.argtypes = [vp, f64, f64]
    lib.ncsSRSetOTFMask.argtypes = [vp, f64]
    lib.ncsSRSetU.argtypes = [vp, f64]
    lib.ncsSRGetU.argtypes = [vp, f64]
    lib.ncsSRSolve.argtypes = [vp, ctypes.c_double, ctypes.c_int]
    lib.ncsSRSolve.restype = ctypes.c_int
    lib.ncsSRCalcLogLikelihood.argtypes = [vp]
    lib.ncsSRCalcLogLikelihood.restype = ctypes.c_double
    lib.ncsSRCalcNoiseContribution.argtypes = [vp]
    lib.ncsSRCalcNoiseContribution.restype = ctypes.c_double
    lib.ncsSRCalcLLGradient.argtypes = [vp, f64]
    lib.ncsSRCalcNCGradient.argtypes = [vp, f64]
    lib.ncso_sr_set_param.argtypes = [vp, ctypes.POINTER(NCSOParam)]
    lib.ncso_sr_last_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_double)]
    lib.ncso_sr_solve_from_u.argtypes = [vp, ctypes.c_double]
    lib.ncso_sr_solve_from_u.restype = ctypes.c_int
    lib.ncso_cost_grad.argtypes = [f64, f64, f64, f64, ctypes.c_double,
                                   ctypes.c_int, vp]
    lib.ncso_cost_grad.restype = ctypes.c_double
    lib.ncso_solve_tiles.argtypes = [f64, f64, vp, f64, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(NCSOParam), ctypes.c_int,
                                     vp, vp, vp, vp, vp]
    lib.ncso_solve_tiles.restype = ctypes.c_int
    lib.ncso_solve_tiles_f32.argtypes = [f32, f32, vp, f32, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(NCSOParam), ctypes.c_int,
                                         vp, vp, vp, vp, vp]
    lib.ncso_solve_tiles_f32.restype = ctypes.c_int
    lib.ncso_solve_tiles_masked_f32.argtypes = [f32, f32, f32, f32, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(NCSOParam), ctypes.c_int,
                                                vp, vp, vp, vp]
    lib.ncso_solve_tiles_masked_f32.restype = ctypes.c_int
    lib.ncso_eval_cost_f32.argtypes = [f32, f32, f32, f32, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       vp, vp, vp]
    lib.ncso_reduce_noise.argtypes = [f64, f64, f64, f64, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(NCSOParam), ctypes.c_int,
                                      vp, vp, vp]
    lib.ncso_reduce_noise.restype = ctypes.c_int
    lib.ncso_tiles_per_axis.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.ncso_tiles_per_axis.restype = ctypes.c_int
    _lib = lib
    return lib


def default_param(**overrides):
    """liblbfgs-1.10 defaults (m=6, eps=1e-5, More-Thuente, unbounded)."""
    p = NCSOParam()
    load().ncso_param_init(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def cost_grad(u, data, gamma, mask_fft, alpha):
    """FP64 cost and gradient of one r x r tile (mask in FFT order)."""
    lib = load()
    u = numpy.ascontiguousarray(u, dtype=numpy.float64)
    r = u.shape[0]
    g = numpy.zeros(u.size, dtype=numpy.float64)
    f = lib.ncso_cost_grad(u, numpy.ascontiguousarray(data, numpy.float64),
                           numpy.ascontiguousarray(gamma, numpy.float64),
                           numpy.ascontiguousarray(mask_fft, numpy.float64),
                           float(alpha), r, _ptr(g))
    return f, g.reshape(u.shape)


def solve_tiles(data, gamma, mask_fft, alpha, u0=None, param=None,
                nthreads=0, f32=False):
    """Solve a stack of independent tiles [ntiles, r, r].

    Returns dict(u, cost, status, iters, evals).  ``f32=True`` passes float32
    tiles straight through (inputs are converted exactly to double inside).
    """
    lib = load()
    dt = numpy.float32 if f32 else numpy.float64
    data = numpy.ascontiguousarray(data, dtype=dt)
    gamma = numpy.ascontiguousarray(gamma, dtype=dt)
    mask = numpy.ascontiguousarray(mask_fft, dtype=dt)
    nt, r = data.shape[0], data.shape[1]
    u = numpy.zeros_like(data)
    cost = numpy.zeros(nt, numpy.float64)
    status = numpy.zeros(nt, numpy.int32)
    iters = numpy.zeros(nt, numpy.int32)
    evals = numpy.zeros(nt, numpy.int32)
    u0p = None if u0 is None else numpy.ascontiguousarray(u0, dtype=dt)
    prm = param if param is not None else default_param()
    fn = lib.ncso_solve_tiles_f32 if f32 else lib.ncso_solve_tiles
    rc = fn(data, gamma, _ptr(u0p), mask, float(alpha), r, nt,
            ctypes.byref(prm), int(nthreads), _ptr(u), _ptr(cost),
            _ptr(status), _ptr(iters), _ptr(evals))
    if rc != 0:
        raise ValueError("oracle solve_tiles rejected its arguments")
    return dict(u=u, cost=cost, status=status, iters=iters, evals=evals)


def solve_tiles_masked(data, gamma, lin, mask_fft, alpha, param=None, nthreads=0):
    """float32 tiles with a 0/1 likelihood weight per pixel (block tiling)."""
    lib = load()
    data = numpy.ascontiguousarray(data, numpy.float32)
    nt, r = data.shape[0], data.shape[1]
    u = numpy.zeros_like(data)
    cost = numpy.zeros(nt, numpy.float64)
    status = numpy.zeros(nt, numpy.int32)
    iters = numpy.zeros(nt, numpy.int32)
    prm = param if param is not None else default_param()
    rc = lib.ncso_solve_tiles_masked_f32(
        data, numpy.ascontiguousarray(gamma, numpy.float32),
        numpy.ascontiguousarray(lin, numpy.float32),
        numpy.ascontiguousarray(mask_fft, numpy.float32), float(alpha), r, nt,
        ctypes.byref(prm), int(nthreads), _ptr(u), _ptr(cost), _ptr(status), _ptr(iters))
    if rc != 0:
        raise ValueError("oracle solve_tiles_masked rejected its arguments")
    return dict(u=u, cost=cost, status=status, iters=iters)


def eval_cost_f32(u, data, gamma, mask_fft, alpha, nthreads=0):
    """FP64 cost, gradient norm and |u| of float32 tiles [ntiles, r, r]."""
    lib = load()
    u = numpy.ascontiguousarray(u, numpy.float32)
    nt, r = u.shape[0], u.shape[1]
    cost = numpy.zeros(nt, numpy.float64)
    gn = numpy.zeros(nt, numpy.float64)
    un = numpy.zeros(nt, numpy.float64)
    lib.ncso_eval_cost_f32(u, numpy.ascontiguousarray(data, numpy.float32),
                           numpy.ascontiguousarray(gamma, numpy.float32),
                           numpy.ascontiguousarray(mask_fft, numpy.float32),
                           float(alpha), r, nt, int(nthreads), _ptr(cost),
                           _ptr(gn), _ptr(un))
    return cost, gn, un


def reduce_noise(image, gamma, mask_fft, alpha, param=None, nthreads=0):
    """Whole-image NCS with the reference tiling; returns (image, tile info)."""
    lib = load()
    image = numpy.ascontiguousarray(image, numpy.float64)
    gamma = numpy.ascontiguousarray(gamma, numpy.float64)
    mask = numpy.ascontiguousarray(mask_fft, numpy.float64)
    r = mask.shape[0]
    nx = lib.ncso_tiles_per_axis(image.shape[0], r)
    ny = lib.ncso_tiles_per_axis(image.shape[1], r)
    out = numpy.zeros_like(image)
    cost = numpy.zeros(max(nx * ny, 0), numpy.float64)
    status = numpy.zeros(max(nx * ny, 0), numpy.int32)
    iters = numpy.zeros(max(nx * ny, 0), numpy.int32)
    prm = param if param is not None else default_param()
    nt = lib.ncso_reduce_noise(out, image, gamma, mask, float(alpha),
                               image.shape[0], image.shape[1], r,
                               ctypes.byref(prm), int(nthreads), _ptr(cost),
                               _ptr(status), _ptr(iters))
    if nt < 0:
        raise ValueError("oracle reduce_noise rejected its arguments")
    return out, dict(cost=cost, status=status, iters=iters, nx=nx, ny=ny)


# ---------------------------------------------------------------------------
# pyCNCS mirror (clib/pyCNCS/ncs_c.py)
# ---------------------------------------------------------------------------

class NCSCException(Exception):
    pass


def checkOTFMask(otf_mask):
    """ncs_c.py:239-247: max(imag(ifft2(fftshift(M)))) <= 1e-12."""
    otf_mask_fft = numpy.fft.ifft2(numpy.fft.fftshift(otf_mask))
    return not (numpy.max(numpy.imag(otf_mask_fft)) > 1.0e-12)


class NCSCSubRegion(object):
    """ncs_c.py:67-236 over the oracle library (centred mask in setOTFMask)."""

    def __init__(self, r_size=None, strict=True, param=None):
        if (r_size % 2) != 0:
            raise NCSCException("Sub region size must be divisible by 2!")
        self.lib = load()
        self.r_size = r_size
        self.strict = strict
        self.c_ncs = self.lib.ncsSRInitialize(r_size)
        if param is not None:
            self.lib.ncso_sr_set_param(self.c_ncs, ctypes.byref(param))

    def calcLLGradient(self):
        g = numpy.zeros(self.r_size * self.r_size)
        self.lib.ncsSRCalcLLGradient(self.c_ncs, g)
        return g

    def calcLogLikelihood(self):
        return self.lib.ncsSRCalcLogLikelihood(self.c_ncs)

    def calcNCGradient(self):
        g = numpy.zeros(self.r_size * self.r_size)
        self.lib.ncsSRCalcNCGradient(self.c_ncs, g)
        return g

    def calcNoiseContribution(self):
        return self.lib.ncsSRCalcNoiseContribution(self.c_ncs)

    def cleanup(self):
        if self.c_ncs is not None:
            self.lib.ncsSRCleanup(self.c_ncs)
            self.c_ncs = None

    def cSolve(self, alpha, verbose=False):
        ret = self.lib.ncsSRSolve(self.c_ncs, alpha, 0)
        self.last_status = ret
        if self.strict and ret != 0:
            raise NCSCException("Solver failed with error code {0:d}!".format(ret))
        return self.getU()

    def lastStats(self):
        it, ev, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self.lib.ncso_sr_last_stats(self.c_ncs, ctypes.byref(it),
                                    ctypes.byref(ev), ctypes.byref(f))
        return it.value, ev.value, f.value

    def getU(self):
        u = numpy.zeros((self.r_size, self.r_size))
        self.lib.ncsSRGetU(self.c_ncs, u)
        return u

    def newRegion(self, image, gamma):
        if self.strict:
            if image.shape[0] != image.shape[1] or image.size != self.r_size ** 2:
                raise NCSCException("Sub-region image size must match sub-region size!")
            if gamma.shape[0] != gamma.shape[1] or gamma.size != self.r_size ** 2:
                raise NCSCException("Sub-region gamma size must match sub-region size!")
        self.image = image
        self.lib.ncsSRNewRegion(self.c_ncs,
                                numpy.ascontiguousarray(image, numpy.float64),
                                numpy.ascontiguousarray(gamma, numpy.float64))

    def setOTFMask(self, otf_mask):
        if self.strict:
            if otf_mask.shape[0] != otf_mask.shape[1]:
                raise NCSCException("OTF must be square!")
            if otf_mask.size != self.r_size * self.r_size:
                raise NCSCException("OTF size must match sub-region size!")
            if not checkOTFMask(otf_mask):
                raise NCSCException("OTF does not have the expected symmetry!")
        tmp = numpy.fft.fftshift(otf_mask)
        self.lib.ncsSRSetOTFMask(self.c_ncs, numpy.ascontiguousarray(tmp, numpy.float64))

    def setU(self, u):
        self.lib.ncsSRSetU(self.c_ncs, numpy.ascontiguousarray(u, numpy.float64))


def cReduceNoise(image, gamma, otf_mask, alpha, strict=True):
    """ncs_c.py:250-276 (FFT-order mask).  Returns float64 for any input
    dtype (the reference returns zeros for non-float64 input, a quirk this
    mirror does not keep)."""
    if strict:
        if otf_mask.shape[0] != otf_mask.shape[1]:
            raise NCSCException("OTF must be square!")
        if (otf_mask.shape[0] % 2) != 0:
            raise NCSCException("Sub region size must be divisible by 2!")
    image = numpy.ascontiguousarray(image, numpy.float64)
    ncs_image = numpy.zeros_like(image)
    load().ncsReduceNoise(ncs_image, image,
                          numpy.ascontiguousarray(gamma, numpy.float64),
                          numpy.ascontiguousarray(otf_mask, numpy.float64),
                          float(alpha), image.shape[0], image.shape[1],
                          otf_mask.shape[0])
    return ncs_image


def pyReduceNoise(image, gamma, otf_mask, alpha, strict=True):
    """ncs_c.py:279-322: Python tiling over the per-tile solver (centred mask)."""
    r_size = otf_mask.shape[0]
    s_size = r_size - 2
    ncs_sr = NCSCSubRegion(r_size=r_size, strict=False)
    ncs_sr.setOTFMask(otf_mask)
    pad_image = numpy.pad(image, 1, "edge")
    pad_gamma = numpy.pad(gamma, 1, "edge")
    ncs_image = numpy.zeros(image.shape, dtype=numpy.float64)
    for i in range(0, pad_image.shape[0], s_size):
        bx = pad_image.shape[0] - r_size if (i + r_size) > pad_image.shape[0] else i
        ex = bx + r_size
        for j in range(0, pad_image.shape[1], s_size):
            by = pad_image.shape[1] - r_size if (j + r_size) > pad_image.shape[1] else j
            ey = by + r_size
            ncs_sr.newRegion(pad_image[bx:ex, by:ey], pad_gamma[bx:ex, by:ey])
            im = ncs_sr.cSolve(alpha)
            ncs_image[bx:ex - 2, by:ey - 2] = im[1:-1, 1:-1]
    ncs_sr.cleanup()
    return ncs_image
