"""Python view of one libmia_hip context (include/mia_hip.h).

Images are numpy arrays in the reference's layout, x fastest: shape (ny, nx)
in 2D and (nz, ny, nx) in 3D.  Parameter vectors are float64, interleaved per
coefficient, coefficients x fastest (mia/3d/transform/spline.cc:289-358).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import GRAD_EXACT, GRAD_REFERENCE, check

__all__ = ["HipContext", "spline_csize", "GRAD_REFERENCE", "GRAD_EXACT"]


def spline_csize(n, rate, dim):
    """Coefficient grid size along one axis (includes the enlargement by 2);
    raises ValueError where the reference throws invalid_argument."""
    v = _lib.load().mia_hip_spline_csize(int(n), float(rate), int(dim))
    if v < 0:
        raise ValueError("spline rate %g is too large for image size %d" % (rate, n))
    return v


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class HipContext:
    """One device context holding one batch of registration instances."""

    def __init__(self, device=0):
        self._L = _lib.load()
        self._h = ctypes.c_void_p()
        check(self._L.mia_hip_ctx_create(int(device), ctypes.byref(self._h)))
        self.dim = None
        self.n_inst = 0
        self.shapes = []
        self.csizes = []
        self._offsets = []
        self.n_params_total = 0

    # -- set-up ---------------------------------------------------------------
    def set_warp_variant(self, variant):
        """Force a warp tile variant (3D or 2D by the setup's dim) for the next
        setup (-1: tuned)."""
        check(self._L.mia_hip_set_warp_variant(self._h, int(variant)))

    def set_stage1_path(self, path):
        """Force the 3D reference-gradient stage-1 layout for the next setup:
        0 per slice, 1 per z knot cell where allowed, -1 tuned."""
        check(self._L.mia_hip_set_stage1_path(self._h, int(path)))

    def warp_variant(self, inst):
        v = ctypes.c_int()
        check(self._L.mia_hip_get_warp_variant(self._h, int(inst), ctypes.byref(v)))
        return v.value

    def setup(self, dim, shapes, csizes, degree=3):
        """shapes: numpy shapes per instance; csizes: coefficient grid sizes per
        instance in (x, y[, z]) order."""
        n = len(shapes)
        sz = np.zeros(3 * n, dtype=np.uint32)
        cs = np.zeros(3 * n, dtype=np.uint32)
        for i, (s, c) in enumerate(zip(shapes, csizes)):
            if len(s) != dim or len(c) != dim:
                raise ValueError("instance %d: dimension mismatch" % i)
            xyz = list(s[::-1]) + [1] * (3 - dim)
            sz[3 * i:3 * i + 3] = xyz
            cs[3 * i:3 * i + 3] = list(c) + [1] * (3 - dim)
        check(self._L.mia_hip_setup(self._h, dim, n, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                    cs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(degree)))
        self.dim = dim
        self.n_inst = n
        self.shapes = [tuple(s) for s in shapes]
        self.csizes = [tuple(c) for c in csizes]
        self._offsets = []
        for i in range(n):
            a, b = ctypes.c_size_t(), ctypes.c_size_t()
            check(self._L.mia_hip_n_params(self._h, i, ctypes.byref(a), ctypes.byref(b)))
            self._offsets.append((b.value, a.value))
        self.n_params_total = sum(n for _, n in self._offsets)

    def setup_grid(self, csizes):
        """New coefficient grids (x, y[, z] per instance) for the instances of
        the last setup, keeping the images already set."""
        if self.n_inst and len(csizes) != self.n_inst:
            raise ValueError("expected %d coefficient grids" % self.n_inst)
        cs = np.zeros(3 * max(self.n_inst, 1), dtype=np.uint32)
        for i, c in enumerate(csizes[:self.n_inst]):
            if len(c) != self.dim:
                raise ValueError("instance %d: dimension mismatch" % i)
            cs[3 * i:3 * i + 3] = list(c) + [1] * (3 - self.dim)
        check(self._L.mia_hip_setup_grid(self._h, cs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        self.csizes = [tuple(c) for c in csizes]
        self._offsets = []
        for i in range(self.n_inst):
            a, b = ctypes.c_size_t(), ctypes.c_size_t()
            check(self._L.mia_hip_n_params(self._h, i, ctypes.byref(a), ctypes.byref(b)))
            self._offsets.append((b.value, a.value))
        self.n_params_total = sum(n for _, n in self._offsets)

    def stage1_path(self, inst):
        """1: reference-gradient stage 1 per (z knot cell, y band); 0: per (slice, y band)."""
        v = ctypes.c_int()
        check(self._L.mia_hip_get_stage1_path(self._h, int(inst), ctypes.byref(v)))
        return v.value

    def set_images_image(self, inst, moving, fixed):
        """Images from two level.DeviceImage objects of the instance's size."""
        check(self._L.mia_hip_set_images_image(self._h, int(inst), moving.handle(), fixed.handle()))

    def set_images(self, inst, moving, fixed):
        moving = np.ascontiguousarray(moving, dtype=np.float32)
        fixed = np.ascontiguousarray(fixed, dtype=np.float32)
        if moving.shape != self.shapes[inst] or fixed.shape != self.shapes[inst]:
            raise RuntimeError("image size mismatch for instance %d" % inst)
        check(self._L.mia_hip_set_images(self._h, int(inst), _f32p(moving), _f32p(fixed)))

    def set_images_device(self, inst, d_moving, d_fixed):
        """Device pointers (ints) of float32 images already in HBM."""
        check(self._L.mia_hip_set_images_device(self._h, int(inst), ctypes.c_void_p(d_moving),
                                                ctypes.c_void_p(d_fixed)))

    def set_penalty(self, inst, penalty, weight):
        """Attach a level.DivcurlPenalty (grid = the instance's) evaluated with
        every evaluation (weight * value added to the cost, -weight * its
        gradient to the gradient), or detach with penalty=None."""
        h = penalty._h if penalty is not None else None
        check(self._L.mia_hip_set_penalty(self._h, int(inst), h, float(weight)))
        self._penalty = penalty  # keep it alive while attached

    def set_cost(self, norm=False, autothresh=0.0, weight=1.0, grad_mode=GRAD_REFERENCE):
        check(self._L.mia_hip_set_cost(self._h, int(bool(norm)), float(autothresh), float(weight), int(grad_mode)))

    # -- evaluation -------------------------------------------------------------
    def param_slice(self, inst):
        off, n = self._offsets[inst]
        return slice(off, off + n)

    def cost_grad(self, params):
        """Returns (costs[n_inst], grad[n_params_total])."""
        params = np.ascontiguousarray(params, dtype=np.float64)
        if self.n_inst == 0:   # let the library report the call-order error
            check(self._L.mia_hip_cost_grad(self._h, None, None, None))
        if params.size != self.n_params_total:
            raise ValueError("expected %d parameters, got %d" % (self.n_params_total, params.size))
        grad = np.empty_like(params)
        costs = np.empty(self.n_inst)
        check(self._L.mia_hip_cost_grad(self._h, _f64p(params), _f64p(grad), _f64p(costs)))
        return costs, grad

    def cost_grad_device(self, d_params, d_grad, d_costs):
        check(self._L.mia_hip_cost_grad_device(self._h, ctypes.c_void_p(d_params), ctypes.c_void_p(d_grad),
                                               ctypes.c_void_p(d_costs)))

    def device_buffers(self):
        """(d_params, d_grad, d_costs) device pointers owned by the context."""
        a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(self._L.mia_hip_device_buffers(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def warp(self, inst, params_inst):
        params_inst = np.ascontiguousarray(params_inst, dtype=np.float64)
        if params_inst.size != self._offsets[inst][1]:
            raise ValueError("wrong parameter count for instance %d" % inst)
        out = np.empty(self.shapes[inst], dtype=np.float32)
        check(self._L.mia_hip_warp(self._h, int(inst), _f64p(params_inst), _f32p(out)))
        return out

    # -- graph ------------------------------------------------------------------
    def graph_capture(self):
        check(self._L.mia_hip_graph_capture(self._h))
        p = ctypes.POINTER(ctypes.c_double)()
        check(self._L.mia_hip_graph_params(self._h, ctypes.byref(p)))
        g = ctypes.POINTER(ctypes.c_double)()
        c = ctypes.POINTER(ctypes.c_double)()
        check(self._L.mia_hip_graph_outputs(self._h, ctypes.byref(g), ctypes.byref(c)))
        self._gp = np.ctypeslib.as_array(p, shape=(self.n_params_total,))
        self._gg = np.ctypeslib.as_array(g, shape=(self.n_params_total,))
        self._gc = np.ctypeslib.as_array(c, shape=(self.n_inst,))

    def graph_run(self, params):
        """Replay the captured iteration; returns views of the pinned outputs
        (valid until the next replay)."""
        self._gp[:] = params
        check(self._L.mia_hip_graph_run(self._h))
        return self._gc, self._gg

    # -- on-device optimizer --------------------------------------------------------
    def optimize(self, x0, maxeval=100, memory=10, first_step=1.0, iters_per_graph=10):
        """On-device L-BFGS for all instances; returns (x, costs, evals, seconds)."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        if x0.size != self.n_params_total:
            raise ValueError("expected %d parameters, got %d" % (self.n_params_total, x0.size))
        x = np.empty_like(x0)
        costs = np.empty(self.n_inst)
        evals = np.empty(self.n_inst, dtype=np.int32)
        sec = ctypes.c_double()
        check(self._L.mia_hip_optimize(self._h, _f64p(x0), int(maxeval), int(memory), float(first_step),
                                       int(iters_per_graph), x.ctypes.data, costs.ctypes.data, evals.ctypes.data,
                                       ctypes.byref(sec)))
        return x, costs, evals, sec.value

    @staticmethod
    def optimize_many(contexts, x0s, maxeval=100, memory=10, first_step=1.0, iters_per_graph=10):
        """On-device L-BFGS on several contexts of one device at once, their
        iterations overlapping (mia_hip_optimize_many); returns one
        (x, costs, evals) per context and the wall time of the joint loop."""
        if len(contexts) != len(x0s) or not contexts:
            raise ValueError("need one starting point per context")
        L = contexts[0]._L
        n = len(contexts)
        x0s = [np.ascontiguousarray(x, dtype=np.float64) for x in x0s]
        for c, x in zip(contexts, x0s):
            if x.size != c.n_params_total:
                raise ValueError("expected %d parameters, got %d" % (c.n_params_total, x.size))
        xs = [np.empty_like(x) for x in x0s]
        costs = [np.empty(c.n_inst) for c in contexts]
        evals = [np.empty(c.n_inst, dtype=np.int32) for c in contexts]
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        hs = (ctypes.c_void_p * n)(*[c._h.value for c in contexts])
        ax0 = (dp * n)(*[_f64p(x) for x in x0s])
        ax = (dp * n)(*[_f64p(x) for x in xs])
        ac = (dp * n)(*[_f64p(x) for x in costs])
        ae = (ip * n)(*[e.ctypes.data_as(ip) for e in evals])
        sec = ctypes.c_double()
        check(L.mia_hip_optimize_many(hs, n, ax0, int(maxeval), int(memory), float(first_step), int(iters_per_graph),
                                      ax, ac, ae, ctypes.byref(sec)))
        return [(xs[k], costs[k], evals[k]) for k in range(n)], sec.value

    def set_optimize_launch(self, mode):
        """How optimize() queues its chunks: "stream" (direct launches, the default) or "graph"."""
        modes = {"stream": 0, "graph": 1}
        if mode not in modes:
            raise ValueError("optimize launch mode must be 'stream' or 'graph', got %r" % (mode,))
        check(self._L.mia_hip_set_optimize_launch(self._h, modes[mode]))

    # -- misc ---------------------------------------------------------------------
    def synchronize(self):
        check(self._L.mia_hip_synchronize(self._h))

    def stream(self):
        s = ctypes.c_void_p()
        check(self._L.mia_hip_get_stream(self._h, ctypes.byref(s)))
        return s.value

    def profile_enable(self, on=True):
        check(self._L.mia_hip_profile_enable(self._h, int(bool(on))))

    def profile_read(self):
        ms = (ctypes.c_double * 16)()
        n = (ctypes.c_int * 16)()
        ns = ctypes.c_int()
        check(self._L.mia_hip_profile_read(self._h, ms, n, 16, ctypes.byref(ns)))
        return {self._L.mia_hip_stage_name(k).decode(): (ms[k], n[k]) for k in range(ns.value)}

    def close(self):
        if self._h:
            self._L.mia_hip_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
