"""OTF masks for NCS (16 x 16, centred unless stated otherwise).

Every mask here is real and point-symmetric on the centred grid, so it passes
the reference's symmetry check (clib/pyCNCS/ncs_c.py:239-247).  Masks are
returned CENTRED (DC at index [r//2, r//2]), the convention of
reduceNoise (opencl/pyOpenCLNCS/ncs.py:208-222); ``to_fft_order`` gives the
order the solvers consume.
"""
import numpy


def to_fft_order(mask):
    """Centred -> FFT order (numpy.fft.fftshift, as ncs.py:222 / ncs_c.py:220)."""
    return numpy.fft.fftshift(mask)


def kspace(r, pixel_size):
    """Radial spatial frequency (1/um) on the centred r x r grid
    (python3-6/pyNCS/denoisetools.py:96-101)."""
    X, Y = numpy.meshgrid(numpy.arange(-r / 2, r / 2, 1), numpy.arange(-r / 2, r / 2, 1))
    return numpy.sqrt(X ** 2 + Y ** 2) / (r * pixel_size)


def otf_mask(r=16, pixel_size=0.1, NA=1.4, wavelength=0.7):
    """1 - incoherent circular-pupil OTF with cutoff 2NA/lambda.

    T(rho) = (2/pi)(acos rho - rho sqrt(1 - rho^2)), rho = k / (2 NA / lambda),
    T = 0 beyond rho = 1; the mask is M = 1 - T (DC -> 0, beyond the
    cutoff -> 1).  Pixel size and wavelength in um.
    """
    rho = kspace(r, pixel_size) / (2.0 * NA / wavelength)
    rho_c = numpy.minimum(rho, 1.0)
    T = (2.0 / numpy.pi) * (numpy.arccos(rho_c) - rho_c * numpy.sqrt(1.0 - rho_c ** 2))
    T[rho >= 1.0] = 0.0
    return 1.0 - T


def psf_otf(r=16, pixel_size=0.1, NA=1.4, wavelength=0.7):
    """PSF, pupil and OTF of python3-6/pyNCS/denoisetools.py:103-115
    (genpsfparam), centred r x r.

    The pupil takes the value of the radial frequency |k| inside the
    aperture |k| <= NA / lambda and 0 outside (the reference assigns the
    frequency grid itself to the pupil, then zeroes it past the cutoff).
    psfA = fftshift(fft2(pupil)), PSF = |psfA|^2 normalised to unit sum,
    OTF = |fftshift(ifft2(|psfA|^2))| normalised to a maximum of 1: the
    pupil's circular autocorrelation, with support up to 2 NA / lambda.
    """
    pupil = kspace(r, pixel_size)
    pupil[pupil > NA / wavelength] = 0.0
    psf_a = numpy.fft.fftshift(numpy.fft.fft2(pupil))
    psf = (psf_a * psf_a.conj()).real
    otf = numpy.abs(numpy.fft.fftshift(numpy.fft.ifft2(psf_a * psf_a.conj())))
    return psf / psf.sum(), pupil, otf / otf.max()


def psf_otf_mask(r=16, pixel_size=0.1, NA=1.4, wavelength=0.7):
    """1 - the PSF-derived OTF of psf_otf (centred): 0 at DC, 1 past the
    OTF's cutoff 2 NA / lambda."""
    return 1.0 - psf_otf(r, pixel_size, NA, wavelength)[2]


def gaussian_highpass_mask(r=16):
    """The Gaussian high-pass mask of opencl/pyOpenCLNCS/py_ref.py:129-137,
    returned centred (that function returns it in FFT order)."""
    x = (numpy.arange(r) - r / 2.0) / 3.0
    gx = numpy.exp(-x * x)
    m = 1.0 - numpy.outer(gx, gx)
    m = m - numpy.min(m)
    return m / numpy.max(m)


def random_mask(r, rng):
    """Real part of the FFT of uniform noise rescaled to [0, 1]
    (clib/pyCNCS/test/py_ref.py:77-85).  Point-symmetric but not even in
    each axis separately: it exercises the general mask path."""
    psf = rng.uniform(0.0, 1.0, size=(r, r))
    m = numpy.real(numpy.fft.fftshift(numpy.fft.fft2(psf)))
    m = m - numpy.min(m)
    return m / numpy.max(m)


def raised_cosine_mask(r=16, pixel_size=0.1, NA=1.4, wavelength=0.7,
                       kind="OTFweighted", w=1.0, h=0.7):
    """1 - raised-cosine low-pass filter (python3-6/pyNCS/denoisetools.py:227-261).

    kind: "OTFweighted" (beta=1, T=lambda/(4 NA 1.4)), "pureN" (beta=0.2,
    T=(1-beta) lambda / (4 NA)) or "adjustable" (w, h parameters).
    """
    kr = kspace(r, pixel_size)
    kmax = 1.0 / numpy.sqrt(2.0) / pixel_size
    if kind == "pureN":
        beta = 0.2
        T = (1 - beta) * wavelength / 4.0 / NA
    elif kind == "adjustable":
        w0 = w * NA / wavelength
        beta = numpy.pi / 2 * (kmax / w0 - 1) / (numpy.arccos(1 - 2 * h) + numpy.pi / 2 * (kmax / w0 - 1))
        T = (1 - beta) / w0 / 2
    elif kind == "OTFweighted":
        beta = 1.0
        T = wavelength / 4.0 / NA / 1.4
    else:
        raise ValueError("unknown filter kind " + repr(kind))
    f = 0.5 * (1 + numpy.cos(numpy.pi * T / beta * (kr - (1 - beta) / 2 / T)))
    f[kr < (1 - beta) / 2 / T] = 1.0
    f[kr > (1 + beta) / 2 / T] = 0.0
    return 1.0 - f
