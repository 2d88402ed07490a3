"""CPU-side checks of the product boundary: the C-ABI library loads and exports
every entry point include/mia_hip.h declares, the ctypes table matches the
header, and the pure host functions agree with the oracle.  No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mia_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from mia_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared(_lib.HEADER)
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_ctypes_table_matches_header():
    from mia_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared(_lib.HEADER)


def test_optimize_many_rejects_bad_lists():
    """mia_hip_optimize_many checks its context list before touching a device:
    no contexts, a null list or a null context are argument errors."""
    from mia_amd import _lib
    L = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    x0 = (dp * 1)()
    hs = (ctypes.c_void_p * 1)(None)
    assert L.mia_hip_optimize_many(None, 1, x0, 5, 10, 1.0, 10, None, None, None, None) == -1
    assert L.mia_hip_optimize_many(hs, 0, x0, 5, 10, 1.0, 10, None, None, None, None) == -1
    assert L.mia_hip_optimize_many(hs, 1, x0, 5, 10, 1.0, 10, None, None, None, None) == -1
    assert b"null context 0" in L.mia_hip_last_error()


@pytest.mark.parametrize("dim", [2, 3])
def test_spline_csize_matches_oracle(dim):
    import mia_amd
    from mia_amd import _lib
    L = _lib.load()
    for n in (10, 33, 96, 128, 160, 192, 224, 256):
        for rate in (1.0, 2.5, 5.0, 6.0, 7.9, 8.0, 10.0, 16.0, 40.0):
            ref = O.spline_csize(n, rate, dim)
            assert L.mia_hip_spline_csize(n, rate, dim) == ref
            if ref < 0:
                with pytest.raises(ValueError):
                    mia_amd.spline_csize(n, rate, dim)


def test_config_grid_sizes():
    import mia_amd
    assert mia_amd.spline_csize(256, 8, 3) == 34          # 34^3 x 3 = 117,912 parameters
    assert mia_amd.spline_csize(128, 16, 2) == 10         # 10^2 x 2 = 200 parameters
    assert mia_amd.spline_csize(96, 8, 2) == 14 and mia_amd.spline_csize(224, 8, 2) == 30


@pytest.mark.parametrize("shape,rate", [((37, 29), 5.0), ((21, 18, 25), 4.0)])
def test_synth_displacement_matches_oracle(shape, rate):
    from mia_amd import synth
    dim = len(shape)
    cs = tuple(O.spline_csize(n, rate, dim) for n in shape[::-1])
    p = synth.random_spline_params(cs, 1.5, 3.0, 1)
    np.testing.assert_allclose(synth.spline_displacement(shape, cs, p), O.displacement(shape, cs, p),
                               rtol=1e-12, atol=1e-12)


def test_synth_warp_matches_oracle_interior():
    """scipy's order-3 mirror spline warp equals the oracle's warp away from
    the clamp (the synthetic moving images use it)."""
    from mia_amd import synth
    shape = (40, 44)
    img = synth.gaussian_blobs(shape, 4, (5, 9), (0.3, 1.0), 0.0, 3)
    cs = (O.spline_csize(44, 8.0, 2), O.spline_csize(40, 8.0, 2))
    p = synth.random_spline_params(cs, 1.5, 3.0, 2)
    W = synth.spline_warp(img, cs, p)
    Wr = O.warp(img, O.displacement(shape, cs, p), 3)
    inside = (Wr > 0.0) & (Wr < img.max())
    np.testing.assert_allclose(W[inside], Wr[inside], atol=2e-6)


def test_no_oracle_in_product():
    """The product package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "mia_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hh", ".cc", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("the oracle", ""), f
