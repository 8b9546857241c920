"""CPU checks of the mask generators and the synthetic-data helpers.

The OTF mask is pinned by the closed form of the incoherent circular-pupil
OTF at rho = 1/2, T = 2/3 - sqrt(3)/(2 pi) (a textbook value, independent of
the implementation's acos form); the raised-cosine filters of
python3-6/pyNCS/denoisetools.py:227-261 by their defining properties (flat
pass band, flat stop band, 1/2 at the centre of the roll-off, monotone).
"""
import numpy
import pytest

from ncs_amd import masks, ncs, sim


def test_otf_mask_closed_form_and_shape():
    # pixel 0.125 um puts a grid point at rho = 1/2: k = 4 / (16 * 0.125)
    # = 2 / um, cutoff 2 NA / lambda = 4 / um
    m = masks.otf_mask(r=16, pixel_size=0.125, NA=1.4, wavelength=0.7)
    t_half = 2.0 / 3.0 - numpy.sqrt(3.0) / (2.0 * numpy.pi)
    assert abs(m[8, 12] - (1.0 - t_half)) < 1e-12
    assert m[8, 8] == 0.0                    # DC passes the OTF: no noise weight
    assert m[0, 0] == 1.0                    # beyond the cutoff
    assert numpy.allclose(m, m[::-1, ::-1][numpy.r_[15, 0:15]][:, numpy.r_[15, 0:15]])
    kr = masks.kspace(16, 0.125)
    order = numpy.argsort(kr.ravel(), kind="stable")
    assert numpy.all(numpy.diff(m.ravel()[order]) >= -1e-12)   # monotone in |k|
    assert ncs.checkOTFMask(m)


@pytest.mark.parametrize("kind", ["OTFweighted", "pureN", "adjustable"])
def test_raised_cosine_masks(kind):
    m = masks.raised_cosine_mask(kind=kind)
    assert m.shape == (16, 16) and m.min() >= 0.0 and m.max() <= 1.0
    assert m[8, 8] == 0.0
    kr = masks.kspace(16, 0.1)
    order = numpy.argsort(kr.ravel(), kind="stable")
    assert numpy.all(numpy.diff(m.ravel()[order]) >= -1e-12)
    assert ncs.checkOTFMask(m)


def test_raised_cosine_midpoint_is_one_half():
    # OTFweighted: beta = 1, T = lambda / (4 NA 1.4); the roll-off centre
    # kr = 1/(2T) is where the filter is 1/2.  Choose the pixel size that puts
    # a grid point there: kr = 3 / (16 p) = 1/(2T)  ->  p = 6 T / 16.
    lam, na = 0.7, 1.4
    T = lam / 4.0 / na / 1.4
    p = 6.0 * T / 16.0
    m = masks.raised_cosine_mask(r=16, pixel_size=p, NA=na, wavelength=lam)
    assert abs(m[8, 11] - 0.5) < 1e-12


def test_fft_order_and_symmetry_check():
    m = masks.otf_mask()
    f = masks.to_fft_order(m)
    assert f[0, 0] == m[8, 8]
    bad = m.copy()
    bad[8, 9] += 0.5          # breaks point symmetry
    assert ncs.checkOTFMask(m) and not ncs.checkOTFMask(bad)


def test_gaussian_and_random_masks_are_normalised():
    g = masks.gaussian_highpass_mask()
    assert abs(g.min()) < 1e-12 and abs(g.max() - 1.0) < 1e-12
    r = masks.random_mask(16, numpy.random.default_rng(3))
    assert abs(r.min()) < 1e-12 and abs(r.max() - 1.0) < 1e-12
    assert ncs.checkOTFMask(r)


def test_photon_conversion_and_gamma():
    adu = numpy.array([[90.0, 100.0, 104.0]])
    off = numpy.full((1, 3), 100.0)
    gain = numpy.full((1, 3), 2.0)
    d = sim.to_photons(adu, off, gain)
    assert d.tolist() == [[1e-6, 1e-6, 2.0]]      # values <= 0 -> 1e-6
    f = sim.make_frame(5, (32, 40))
    g = f["var"] / f["gain"] ** 2
    assert numpy.allclose(f["gamma"], g.astype(numpy.float32))
    f2 = sim.make_frame(5, (32, 40))
    assert numpy.array_equal(f["data"], f2["data"])     # seeded


@pytest.mark.parametrize("r,pixel", [(16, 0.1), (12, 0.1), (16, 0.08)])
def test_psf_otf_against_direct_sums(r, pixel):
    """genpsfparam (denoisetools.py:103-115) by routes that use no FFT: the
    PSF as |sum_x pupil(x) e^(-2 pi i k x / r)|^2 from an explicit DFT
    matrix, the OTF as the pupil's circular autocorrelation summed over
    shifts; both normalised as the reference does."""
    psf, pupil, otf = masks.psf_otf(r=r, pixel_size=pixel)
    kr = masks.kspace(r, pixel)
    assert numpy.array_equal(pupil, numpy.where(kr > 1.4 / 0.7, 0.0, kr))
    idx = numpy.arange(r)
    F = numpy.exp(-2j * numpy.pi * numpy.outer(idx, idx) / r)
    P = F @ pupil @ F.T
    psf_direct = numpy.fft.fftshift(numpy.abs(P) ** 2)  # centring only reorders
    assert numpy.allclose(psf, psf_direct / psf_direct.sum(), rtol=0, atol=1e-14)
    ac = numpy.zeros((r, r))
    for dy in range(r):
        for dx in range(r):
            ac[dy, dx] = numpy.sum(pupil * numpy.roll(numpy.roll(pupil, -dy, 0), -dx, 1))
    ac = numpy.fft.fftshift(ac)  # zero shift to the centre
    assert numpy.allclose(otf, ac / ac.max(), rtol=0, atol=1e-12)
    assert otf[r // 2, r // 2] == 1.0 and abs(psf.sum() - 1.0) < 1e-12


def test_psf_otf_mask_support_and_symmetry():
    """1 - OTF: 0 at DC, 1 past the autocorrelation's support 2 NA / lambda,
    point-symmetric (passes the reference's checkOTFMask)."""
    m = masks.psf_otf_mask()
    kr = masks.kspace(16, 0.1)
    assert m[8, 8] == 0.0
    assert numpy.allclose(m[kr > 2.0 * 1.4 / 0.7 + 1e-9], 1.0, atol=1e-12)
    assert numpy.all((m >= -1e-12) & (m <= 1.0 + 1e-12))
    assert ncs.checkOTFMask(m)
