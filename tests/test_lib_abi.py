"""CPU checks of the C ABI: the library loads, exports every symbol that
include/ncs_hip.h declares, and its host-only functions behave."""
import ctypes
import os
import re

import numpy
import pytest

from ncs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z_0-9]*)\s*\(", text):
            name = m.group(1)
            if name not in ("if", "sizeof") and (name.startswith("ncs") or name == "ncsReduceNoise"):
                names.append(name)
    return sorted(set(names))


def test_header_declares_the_expected_entry_points():
    names = _declared_functions()
    assert "ncs_hip_reduce_frames" in names and "ncsReduceNoise" in names
    assert len(names) >= 12


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header():
    assert set(_declared_functions()) <= set(_lib.SIGNATURES)


def test_default_params():
    p = _lib.default_params()
    assert p.max_iterations == 200 and p.history == 5
    assert numpy.isclose(p.epsilon, 1e-5) and p.linesearch == _lib.LS_MORE_THUENTE
    assert numpy.isclose(p.ftol, 1e-4) and numpy.isclose(p.gtol, 0.9)
    assert p.precond == 2 and _lib.liblbfgs_params().precond == 0
    assert p.precision == _lib.PRECISION_FP32  # FP32 basis changes unless asked
    k = _lib.kernel_semantics_params()
    assert k.history == 8 and k.linesearch == _lib.LS_ARMIJO and numpy.isclose(k.min_step, 1e-6)


def test_tile_counts_match_the_oracle_tiling():
    import oracle
    lib = oracle.load()
    for h in (14, 15, 16, 27, 28, 29, 30, 100, 255, 256, 257, 2048, 2304):
        for w in (14, 19, 256, 2048):
            assert _lib.tile_count(h, w) == lib.ncso_tiles_per_axis(h, 16) * lib.ncso_tiles_per_axis(w, 16)
    assert _lib.tile_count(2048, 2048) == 147 * 147
    assert _lib.tile_count(2048, 2048, _lib.GEOM_BLOCKS) == 128 * 128
    assert _lib.tile_count(2304, 2304, _lib.GEOM_BLOCKS) == 144 * 144
    assert _lib.tile_count(17, 33, _lib.GEOM_BLOCKS) == 2 * 3
    assert _lib.tile_count(13, 20) == -1
    assert _lib.tile_count(20, 20, 7) == -1


def test_context_on_a_machine_without_gpu_reports_an_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.NCSHipError):
        _lib.Context(0)
    assert "device" in _lib.last_error().lower()


def test_drop_in_rejects_unsupported_subregions_without_touching_memory():
    """Odd sizes (the reference refuses them, ncs.c:54-58), sizes past the
    16 x 16 tile and frames smaller than r - 2 (where the reference would
    write out of bounds) are refused before any device work."""
    lib = _lib.load()
    img = numpy.ones((20, 20))
    for r, h in ((9, 20), (18, 20), (12, 9)):
        out = numpy.full((20, 20), 7.0)
        lib.ncsReduceNoise(out.ctypes.data, img.ctypes.data, img.ctypes.data,
                           numpy.ones((r, r)).ctypes.data, 0.1, h, 20, r)
        assert numpy.all(out == 7.0), r
        err = _lib.last_error()
        assert ("16" in err) if r != 12 else ("bad arguments" in err), (r, err)


def test_tile_counts_follow_the_window_side():
    """ncs.c:71-143 with windows of side r: starts -1, -1 + (r - 2), ...
    until a window ends at or past the padded edge (that one clamped)."""
    def loop(n, r):
        count, i = 0, -1
        while True:
            count += 1
            if i + r >= n + 1:
                return count
            i += r - 2
    for r in (4, 6, 10, 12, 16):
        for n in (r - 2, r - 1, r, 30, 31, 97, 128):
            assert _lib.tile_count(n, n, _lib.GEOM_REFERENCE, r) == loop(n, r) ** 2, (r, n)
        assert _lib.tile_count(r - 3, 40, _lib.GEOM_REFERENCE, r) == -1
    # blocks of side r: ceil(n / r) per axis
    for r in (4, 10, 12, 16):
        for n in (1, r - 1, r, r + 1, 40, 97):
            assert _lib.tile_count(n, 40, _lib.GEOM_BLOCKS, r) == (-(-n // r)) * (-(-40 // r))
    assert _lib.tile_count(0, 40, _lib.GEOM_BLOCKS, 12) == -1
    assert _lib.tile_count(40, 40, _lib.GEOM_BLOCKS, 7) == -1
    assert _lib.tile_count(40, 40, _lib.GEOM_REFERENCE, 7) == -1
