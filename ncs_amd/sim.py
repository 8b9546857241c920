"""Synthetic sCMOS data for the NCS path: WLC filament scenes and camera maps.

The scene and camera models follow the simulation helpers of the reference
(python3-6/pyNCS/denoisetools.py:150-225: ideal image, per-pixel gain,
offset and variance, ADU = gain * Poisson + N(0, var) + offset, conversion
back to photo-electrons with values <= 0 set to 1e-6), with the parameters
of the benchmark workloads:

  * filaments: 2-D worm-like chains, persistence length 1 um, tangent-angle
    random walk with std sqrt(ds / Lp) per ds = 10 nm step, lengths
    U(5, 15) um, uniform start and orientation, 40 filaments per
    25.6 x 25.6 um (scaled by area), 2000 photons per um of filament;
  * PSF: Gaussian, sigma = 0.21 lambda / NA (NA 1.4, lambda 700 nm),
    integrated over 100 nm pixels; background 20 photons / pixel;
  * camera: offset ~ N(100, 2) ADU, gain ~ N(2.0, 0.2) ADU/e-, read
    variance ~ Exp(mean 50) ADU^2 capped at 3000 (optionally 1 % of pixels
    at lognormal(7, 0.5)); gamma = var / gain^2.

Everything is seeded through numpy.random.default_rng and deterministic.
"""
import numpy
from scipy.special import erf

PIXEL_UM = 0.1
NA = 1.4
WAVELENGTH_UM = 0.7
SIGMA_PX = 0.21 * WAVELENGTH_UM / NA / PIXEL_UM  # 1.05 px
FILAMENTS_PER_UM2 = 40.0 / (25.6 * 25.6)


def wlc_filaments(rng, n_filaments, height_um, width_um, lp_um=1.0,
                  ds_um=0.01, length_um=(5.0, 15.0)):
    """Point emitters (row_um, col_um) of n 2-D worm-like chains, one array
    per filament, each emitter standing for ds_um of filament."""
    out = []
    step_std = numpy.sqrt(ds_um / lp_um)
    for _ in range(n_filaments):
        n_seg = max(1, int(round(rng.uniform(*length_um) / ds_um)))
        theta = rng.uniform(0, 2 * numpy.pi) + numpy.cumsum(
            rng.normal(0.0, step_std, n_seg))
        r0 = rng.uniform(0, height_um)
        c0 = rng.uniform(0, width_um)
        rows = r0 + numpy.cumsum(ds_um * numpy.cos(theta))
        cols = c0 + numpy.cumsum(ds_um * numpy.sin(theta))
        out.append(numpy.stack([rows, cols], axis=1))
    return out


def render(points_px, weights, shape, sigma_px=SIGMA_PX, half=3, chunk=262144):
    """Sum of pixel-integrated Gaussians (erf) of the given emitters.

    points_px: (N, 2) positions in pixel units (pixel (i, j) spans
    [i, i+1) x [j, j+1)); weights: (N,) photons per emitter."""
    H, W = shape
    img = numpy.zeros(H * W, dtype=numpy.float64)
    offs = numpy.arange(-half, half + 1)
    k = 1.0 / (numpy.sqrt(2.0) * sigma_px)
    for s in range(0, len(points_px), chunk):
        p = points_px[s:s + chunk]
        w = weights[s:s + chunk]
        base = numpy.floor(p).astype(numpy.int64)
        ix = base[:, 0:1] + offs[None, :]
        iy = base[:, 1:2] + offs[None, :]
        ex = 0.5 * (erf((ix + 1 - p[:, 0:1]) * k) - erf((ix - p[:, 0:1]) * k))
        ey = 0.5 * (erf((iy + 1 - p[:, 1:2]) * k) - erf((iy - p[:, 1:2]) * k))
        vals = (w[:, None, None] * ex[:, :, None] * ey[:, None, :]).reshape(-1)
        ii = numpy.broadcast_to(ix[:, :, None], (len(p), 2 * half + 1, 2 * half + 1)).reshape(-1)
        jj = numpy.broadcast_to(iy[:, None, :], (len(p), 2 * half + 1, 2 * half + 1)).reshape(-1)
        ok = (ii >= 0) & (ii < H) & (jj >= 0) & (jj < W)
        img += numpy.bincount(ii[ok] * W + jj[ok], weights=vals[ok], minlength=H * W)
    return img.reshape(H, W)


def filament_scene(rng, shape, photons_per_um=2000.0, background=20.0,
                   blink_p=None, segment_um=None):
    """Ideal photon image of a WLC filament scene with background.

    blink_p: if set, each filament segment of length segment_um is on with
    probability blink_p (per-frame blinking of the time-series workload)."""
    H, W = shape
    n_fil = max(1, int(round(FILAMENTS_PER_UM2 * H * W * PIXEL_UM ** 2)))
    fils = wlc_filaments(rng, n_fil, H * PIXEL_UM, W * PIXEL_UM)
    pts = numpy.concatenate(fils, axis=0) / PIXEL_UM
    w = numpy.full(len(pts), photons_per_um * 0.01)
    if blink_p is not None:
        seg = max(1, int(round((segment_um or 1.0) / 0.01)))
        nseg = (len(pts) + seg - 1) // seg
        on = rng.random(nseg) < blink_p
        w = w * numpy.repeat(on, seg)[:len(pts)]
    return render(pts, w, shape) + background


def camera_maps(rng, shape, heavy_tail=False):
    """Per-pixel offset (ADU), gain (ADU/e-) and read variance (ADU^2)."""
    offset = rng.normal(100.0, 2.0, shape)
    gain = rng.normal(2.0, 0.2, shape)
    var = numpy.minimum(rng.exponential(50.0, shape), 3000.0)
    if heavy_tail:
        hot = rng.random(shape) < 0.01
        var[hot] = rng.lognormal(7.0, 0.5, int(hot.sum()))
    return offset, gain, var


def expose(rng, ideal, offset, gain, var):
    """ADU frame of an ideal photon image (denoisetools.py:205-210)."""
    return gain * rng.poisson(ideal) + numpy.sqrt(var) * rng.standard_normal(ideal.shape) + offset


def to_photons(adu, offset, gain):
    """(ADU - offset) / gain with values <= 0 set to 1e-6
    (denoisetools.py:218-224)."""
    d = (adu - offset) / gain
    d[d <= 0] = 1e-6
    return d


def make_frame(seed, shape, heavy_tail=False, camera_seed=None):
    """One synthetic frame: dict(data=photons f32, gamma=f32, ideal, adu, ...).

    camera_seed fixes the camera maps independently of the scene seed."""
    rng = numpy.random.default_rng(seed)
    crng = numpy.random.default_rng(camera_seed) if camera_seed is not None else rng
    offset, gain, var = camera_maps(crng, shape, heavy_tail=heavy_tail)
    ideal = filament_scene(rng, shape)
    adu = expose(rng, ideal, offset, gain, var)
    data = to_photons(adu, offset, gain)
    gamma = var / gain ** 2
    return dict(data=data.astype(numpy.float32), gamma=gamma.astype(numpy.float32),
                ideal=ideal, adu=adu, offset=offset, gain=gain, var=var)
