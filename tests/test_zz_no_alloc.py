"""The allocation-free entry point (mia_hip_cost_grad_noalloc) through the C
ABI against the context entry point (mia_hip_setup + mia_hip_set_images +
mia_hip_cost_grad) on the same inputs: the same kernels on the same data, so
cost and gradient must be bitwise equal, and within the parity bounds of the
FP64 oracle (cost 1e-5 relative, gradient 1e-5 of the max-norm).  Several
calls in a row share the one scratch array, each smaller than the one before
it, so a stale element left by a larger call would show.

Device memory for the inputs and outputs comes from the HIP runtime the
library itself links (hipMalloc through ctypes): test plumbing, not the
entry point's.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _hip():
    from mia_amd import _lib
    _lib.load()  # the runtime libmia_hip.so links is loaded: same object by soname
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


class _Dev:
    """Device copies of numpy arrays; freed on exit."""

    def __init__(self, H):
        self.H = H
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.H.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p.value)
        assert self.H.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.H.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def __enter__(self):
        return self

    def __exit__(self, *a):
        for p in self.ptrs:
            self.H.hipFree(p)


def _noalloc(mov, fix, cs, p, degree=3, norm=False, thr=0.0, weight=1.0, mode=0):
    from mia_amd import _lib
    L = _lib.load()
    H = _hip()
    dim = mov.ndim
    size = np.array(list(mov.shape[::-1]), dtype=np.uint32)
    csz = np.array(list(cs), dtype=np.uint32)
    with _Dev(H) as dv:
        dm = dv.put(mov.astype(np.float32))
        df = dv.put(fix.astype(np.float32))
        dp = dv.put(p.astype(np.float64))
        dg = dv.put(np.full(p.size, np.nan))
        dc = dv.put(np.full(1, np.nan))
        st = L.mia_hip_cost_grad_noalloc(0, dim, size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         csz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), degree, int(norm),
                                         float(thr), float(weight), mode, dm, df, dp, dg, dc, None)
        _lib.check(st)
        assert H.hipDeviceSynchronize() == 0
        return dv.get(dc, np.zeros(1))[0], dv.get(dg, p.astype(np.float64))


def _context(mov, fix, cs, p, degree=3, norm=False, thr=0.0, weight=1.0, mode=0):
    import mia_amd
    with mia_amd.HipContext(0) as c:
        c.setup(mov.ndim, [mov.shape], [cs], degree)
        c.set_images(0, mov, fix)
        c.set_cost(norm=norm, autothresh=thr, weight=weight, grad_mode=mode)
        costs, g = c.cost_grad(p)
    return costs[0], g


# largest first: each call's scratch layout is a prefix of the previous one's
CASES = [
    ((72, 60, 84), 6.0, 3, dict()),
    ((40, 36, 44), 8.0, 3, dict(mode=1, thr=0.3)),
    ((33, 29, 31), 5.0, 1, dict(norm=True, weight=0.5)),
    ((150, 170), 8.0, 3, dict()),
    ((64, 90), 16.0, 3, dict(mode=1, norm=True)),
]


def test_noalloc_equals_context_calls_in_a_row():
    from mia_amd import synth
    for k, (shape, rate, degree, kw) in enumerate(CASES):
        print("case", k, shape, rate, degree, kw, flush=True)
        dim = len(shape)
        fix = synth.smoothed_noise(shape, 2.0, 40 + k)
        mov = synth.smoothed_noise(shape, 2.0, 60 + k)
        cs = tuple(O.spline_csize(n, rate, dim) for n in shape[::-1])
        p = synth.random_spline_params(cs, 1.5, 3.0, 80 + k)
        c0, g0 = _noalloc(mov, fix, cs, p, degree, **kw)
        c1, g1 = _context(mov, fix, cs, p, degree, **kw)
        assert c0 == c1, (c0, c1)
        assert np.array_equal(g0, g1)
        c_ref, g_ref = O.cost_grad(mov, fix, cs, p, degree=degree, norm=kw.get("norm", False),
                                   autothresh=kw.get("thr", 0.0), grad_mode=kw.get("mode", 0),
                                   weight=kw.get("weight", 1.0))
        assert c0 == pytest.approx(c_ref, rel=1e-5)
        assert np.abs(g0 - g_ref).max() <= 1e-5 * np.abs(g_ref).max()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    import mia_amd
    from mia_amd import _lib
    L = _lib.load()
    size = np.array([256, 256, 256], dtype=np.uint32)   # 16.8 M voxels: needs about 290 MB
    cs = np.array([34, 34, 34], dtype=np.uint32)
    # the size check comes before any launch; the buffers are never touched
    with _Dev(_hip()) as dv:
        b = dv.put(np.zeros(16))
        st = L.mia_hip_cost_grad_noalloc(0, 3, size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         cs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 3, 0, 0.0, 1.0, 0,
                                         b, b, b, b, b, None)
    assert st == mia_amd.ERR_LIMIT
    assert b"scratch" in L.mia_hip_last_error()
