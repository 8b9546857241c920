"""GPU solves with every mask the reference's tools generate.

The three raised-cosine filters of genfilter (python3-6/pyNCS/denoisetools.py:
227-261: OTFweighted, pureN, adjustable with w = 1, h = 0.7 as the notebooks
call it), the PSF-derived OTF of genpsfparam (denoisetools.py:103-115) and
the circular-pupil OTF mask of the benchmark, each on a synthetic filament
frame through reduceNoise: every tile's final cost within 1e-5 relative of
the FP64 oracle's, and the corrected image within 1e-3 of the frame's
largest value of the oracle's (the stop rule |g| <= 1e-5 max(1, |u|) leaves
bright pixels, whose likelihood curvature is small, that much room).  Also
the FP16-split precision on the same inputs.
"""
import numpy
import pytest

import oracle
from ncs_amd import _lib, masks, ncs, sim

pytestmark = pytest.mark.gpu

MASKS = {
    "OTFweighted": lambda: masks.raised_cosine_mask(kind="OTFweighted"),
    "pureN": lambda: masks.raised_cosine_mask(kind="pureN"),
    "adjustable": lambda: masks.raised_cosine_mask(kind="adjustable", w=1.0, h=0.7),
    "psf_otf": lambda: masks.psf_otf_mask(),
    "circular_otf": lambda: masks.otf_mask(),
}


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP32, _lib.PRECISION_F16_SPLIT])
@pytest.mark.parametrize("kind", sorted(MASKS))
def test_generated_masks_match_the_oracle(gpu_ctx, kind, precision):
    print("mask", kind, "precision", precision, flush=True)
    mask = MASKS[kind]()
    assert ncs.checkOTFMask(mask)
    fr = sim.make_frame(300 + len(kind), (128, 160))
    out, info = ncs.reduceNoise([fr["data"]], fr["gamma"], mask, 0.1, return_info=True,
                                params=_lib.default_params(precision=precision))
    ref, rinfo = oracle.reduce_noise(fr["data"].astype(numpy.float64),
                                     fr["gamma"].astype(numpy.float64),
                                     masks.to_fft_order(mask), 0.1)
    rel = numpy.abs(info["cost"].astype(numpy.float64) - rinfo["cost"]) / numpy.abs(rinfo["cost"])
    diff = float(numpy.max(numpy.abs(out[0] - ref)))
    print("  tiles %d, status %s, cost max rel %.2e, pixel max |d| %.2e"
          % (len(rel), numpy.unique(info["status"]).tolist(), rel.max(), diff), flush=True)
    assert numpy.all(info["status"] == 0)
    assert rel.max() <= 1e-5
    assert diff <= 1e-3 * float(numpy.max(numpy.abs(ref)))
