"""ncs_amd -- NCS sCMOS noise correction on AMD Instinct MI355X (gfx950).

The hot path (tiling, per-tile L-BFGS on the NCS cost, stitching) runs in
hand-written HIP (libncs_hip.so, include/ncs_hip.h); this package is its
host side: the reduceNoise drop-in of the reference's Python hosts, the
ctypes binding, synthetic data, OTF masks and the learned warm start.
"""
from .ncs import (NCSHip, NCSHipException, checkOTFMask, cReduceNoise, reduceNoise,
                  solveTiles)

__all__ = ["NCSHip", "NCSHipException", "checkOTFMask", "cReduceNoise", "reduceNoise",
           "solveTiles"]
