"""The bench's roofline reads per-launch SQ counts and HBM traffic from the
rocprofv3 summaries under profiles/, stamped with the hash of the solver
sources they were measured on.  Every kernel and path the bench line reports
must have both figures for the sources as they are now; a solver change
without a new profile would leave the bench line's roofline empty."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

# (kernel, path, precision) of the bench line's roofline entries: the
# headline and the paths it is chosen from, the FP16-split key, and the
# uneven batch's linear start, each at the histories (1, 2 and 3) the
# autotuner chooses between in practice
NEEDED = [
    ("solve_kernel<1,false,false,false>", "linear", "fp32"),
    ("solve_kernel<2,false,false,false>", "linear", "fp32"),
    ("solve_kernel<1,false,false,false>", "tuned", "fp32"),
    ("solve_kernel<2,false,false,false>", "tuned", "fp32"),
    ("solve_kernel<5,false,false,false>", "plain", "fp32"),
    ("solve_kernel<1,false,false,true>", "linear", "f16split"),
    ("solve_kernel<2,false,false,true>", "linear", "f16split"),
    ("solve_kernel<3,false,false,false>", "linear", "fp32"),
    ("solve_kernel<3,false,false,false>", "tuned", "fp32"),
    ("solve_kernel<3,false,false,true>", "linear", "f16split"),
    ("solve_kernel<1,false,true,false>", "uneven_linear", "fp32"),
    ("solve_kernel<2,false,true,false>", "uneven_linear", "fp32"),
    ("solve_kernel<3,false,true,false>", "uneven_linear", "fp32"),
]


@pytest.mark.parametrize("kernel,path,precision", NEEDED)
def test_profile_lines_match_the_solver_sources(kernel, path, precision):
    sq = bench.profiled("sq_per_launch", kernel, path, precision)
    tr = bench.profiled("traffic_per_launch", kernel, path, precision)
    assert sq is not None, ("no SQ profile of %s on %s (%s) for solver sources %s"
                            % (kernel, path, precision, bench.solver_source_hash()))
    assert tr is not None, ("no traffic profile of %s on %s (%s) for solver sources %s"
                            % (kernel, path, precision, bench.solver_source_hash()))
    for key in ("waves", "valu", "mfma", "trans", "salu"):
        assert sq[key] > 0 or (key == "mfma" and precision == "f16split")
    assert sq["valu"] > sq["mfma"]
    assert tr["bytes"] > 0


def test_issue_cycle_model_counts_fp32_mfma_at_32_cycles():
    sq = {"valu": 1000.0, "mfma": 100.0}
    assert bench.issue_cycles(sq, False) == 4 * 900 + 32 * 100
    assert bench.issue_cycles(sq, True) == 4 * 900
