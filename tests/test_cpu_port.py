"""The FP32 CPU port of MIA's data flow (oracle/mia_cpu_port.cc, the bench's
CPU baseline) against the FP64 oracle: 3D and 2D cost + gradient, and the
series baseline's registrations."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _port():
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "lib", "libmia_cpu_port.so"))
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    L.mcp_cost_grad_3d.argtypes = [fp, fp] + [ctypes.c_int] * 6 + [dp, ctypes.c_int, dp, ctypes.c_void_p]
    L.mcp_cost_grad_3d.restype = ctypes.c_double
    L.mcp_cost_grad_2d.argtypes = [fp, fp] + [ctypes.c_int] * 4 + [dp, dp]
    L.mcp_cost_grad_2d.restype = ctypes.c_double
    L.mcp_series_2d.argtypes = [ctypes.c_int, ip, fp, fp, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_void_p]
    L.mcp_series_2d.restype = ctypes.c_double
    return L


def test_port_3d_matches_oracle():
    from mia_amd import synth
    shape = (20, 24, 28)
    fix, mov = synth.smoothed_noise(shape, 2.0, 1), synth.smoothed_noise(shape, 2.0, 2)
    cs = tuple(O.spline_csize(n, 5.0, 3) for n in shape[::-1])
    p = synth.random_spline_params(cs, 1.5, 3.0, 3)
    g = np.zeros_like(p)
    c = _port().mcp_cost_grad_3d(mov, fix, shape[2], shape[1], shape[0], cs[0], cs[1], cs[2], p, 2, g, None)
    c_ref, g_ref = O.cost_grad(mov, fix, cs, p)
    assert c == pytest.approx(c_ref, rel=1e-6)
    assert np.abs(g - g_ref).max() <= 1e-5 * np.abs(g_ref).max()


def test_port_2d_matches_oracle():
    """2D returns the float-accumulated SSD (mia/2d/fullcost/image.cc:99-104)."""
    from mia_amd import synth
    shape = (70, 90)
    fix = synth.gaussian_blobs(shape, seed=3)
    cs = tuple(O.spline_csize(n, 8.0, 2) for n in shape[::-1])
    mov = synth.spline_warp(fix, cs, synth.random_spline_params(cs, 1.5, 3.0, 4)).astype(np.float32)
    p = synth.random_spline_params(cs, 1.0, 2.0, 5)
    g = np.zeros_like(p)
    c = _port().mcp_cost_grad_2d(mov, fix, shape[1], shape[0], cs[0], cs[1], p, g)
    c_ref, g_ref = O.cost_grad(mov, fix, cs, p)
    assert c == pytest.approx(c_ref, rel=1e-4)
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max()


def test_series_baseline_registers():
    from mia_amd import series, synth
    frames = synth.series_frames_2d(4, (48, 80), 8.0, 4.0, seed=9)
    sizes = np.array([v for m, _ in frames for v in m.shape[::-1]], dtype=np.int32)
    movs = np.concatenate([m.ravel() for m, _ in frames]).astype(np.float32)
    fixs = np.concatenate([f.ravel() for _, f in frames]).astype(np.float32)
    costs = np.zeros(4)
    evals = np.zeros(4, dtype=np.int32)
    sec = _port().mcp_series_2d(4, sizes, movs, fixs, 8.0, 30, 2, costs.ctypes.data, evals.ctypes.data)
    assert sec > 0 and np.all(evals == 30)
    for (m, f), c in zip(frames, costs):
        a, b = series.normalize_pair(m, f)
        c0 = 0.5 * float(np.sum((a.astype(np.float64) - b) ** 2))
        assert c < 0.5 * c0
