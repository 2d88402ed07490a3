"""ctypes view of the FP64 oracle (oracle/mia_oracle.c).

Test infrastructure only: imported by tests/, by the smoke check and by the
bench's CPU-baseline leg, always as the checker.  The product path never
imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libmia_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.mo_fastfloor.argtypes = [ctypes.c_double]
        L.mo_fastfloor.restype = ctypes.c_int
        L.mo_bspline3_weights.argtypes = [ctypes.c_double, _dp]
        L.mo_bspline3_dweights.argtypes = [ctypes.c_double, _dp]
        L.mo_bspline3_value.argtypes = [ctypes.c_double]
        L.mo_bspline3_value.restype = ctypes.c_double
        L.mo_mirror_index.argtypes = [ctypes.c_int, ctypes.c_int]
        L.mo_mirror_index.restype = ctypes.c_int
        L.mo_filter_line.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.mo_prefilter.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.mo_interp3.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                 ctypes.c_double, ctypes.c_double, _dp]
        L.mo_interp3.restype = ctypes.c_double
        L.mo_interp2.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                 ctypes.c_double, ctypes.c_double, _dp]
        L.mo_interp2.restype = ctypes.c_double
        L.mo_spline_csize.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int]
        L.mo_spline_csize.restype = ctypes.c_int
        L.mo_spline_scale.argtypes = [ctypes.c_int, ctypes.c_int]
        L.mo_spline_scale.restype = ctypes.c_float
        L.mo_spline_axis_table.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _ip]
        L.mo_spline_derivative_rows.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.mo_displacement.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp]
        L.mo_displacement_at.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.mo_warp.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                              _dp, _dp, ctypes.c_void_p]
        L.mo_image_gradient.argtypes = [ctypes.c_int, _ip, _dp, _dp]
        L.mo_ssd_value.argtypes = [ctypes.c_size_t, _dp, _fp, ctypes.c_int, ctypes.c_float]
        L.mo_ssd_value.restype = ctypes.c_double
        L.mo_ssd_force.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _fp, _dp, ctypes.c_int,
                                   ctypes.c_float, _dp]
        L.mo_ssd_force.restype = ctypes.c_double
        L.mo_translate.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, _dp]
        L.mo_cost_grad.argtypes = [ctypes.c_int, _ip, _fp, _fp, _ip, _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p]
        L.mo_cost_grad.restype = ctypes.c_double
        L.mo_cost.argtypes = [ctypes.c_int, _ip, _fp, _fp, _ip, _dp, ctypes.c_int, ctypes.c_int,
                              ctypes.c_float]
        L.mo_cost.restype = ctypes.c_double
        L.mo_cost_box.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double, _fp, _ip, _dp,
                                  ctypes.c_int, _ip, _ip]
        L.mo_cost_box.restype = ctypes.c_double
        L.mo_gauss_line.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
        L.mo_scale_line.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.mo_scale_image.argtypes = [_fp, _ip, _ip, _fp]
        L.mo_bspline3_deg.argtypes = [ctypes.c_double, ctypes.c_int]
        L.mo_bspline3_deg.restype = ctypes.c_double
        L.mo_integrate2.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double]
        L.mo_integrate2.restype = ctypes.c_double
        L.mo_pp_cells.argtypes = [_ip, _fp, ctypes.c_double, ctypes.c_double, _dp]
        L.mo_divcurl.argtypes = [_ip, _dp, _fp, ctypes.c_void_p]
        L.mo_divcurl.restype = ctypes.c_double
        _lib = L
    return _lib


def _size3(shape):
    """numpy shape (z,y,x) or (y,x) -> int32 [nx, ny, nz]."""
    if len(shape) == 3:
        return np.array([shape[2], shape[1], shape[0]], dtype=np.int32)
    return np.array([shape[1], shape[0], 1], dtype=np.int32)


def fastfloor(v):
    return lib().mo_fastfloor(float(v))


def bspline3_weights(f):
    w = np.zeros(4)
    lib().mo_bspline3_weights(float(f), w)
    return w


def bspline3_dweights(f):
    w = np.zeros(4)
    lib().mo_bspline3_dweights(float(f), w)
    return w


def bspline3_value(x):
    return lib().mo_bspline3_value(float(x))


def mirror_index(i, width):
    return lib().mo_mirror_index(int(i), int(width))


def filter_line(values, degree=3):
    c = np.ascontiguousarray(values, dtype=np.float64).copy()
    lib().mo_filter_line(c, c.size, degree)
    return c


def prefilter(img, degree=3):
    img = np.ascontiguousarray(img, dtype=np.float32)
    s = _size3(img.shape)
    coeff = np.zeros(img.shape, dtype=np.float64)
    vmin, vmax = ctypes.c_double(), ctypes.c_double()
    lib().mo_prefilter(img, int(s[0]), int(s[1]), int(s[2]), degree, coeff, ctypes.byref(vmin),
                       ctypes.byref(vmax))
    return coeff, vmin.value, vmax.value


def interp(coeff, pos, degree=3, vmin=-np.inf, vmax=np.inf, with_grad=False):
    """pos in (x, y[, z]) order."""
    g = np.zeros(3)
    p = np.ascontiguousarray(pos, dtype=np.float64)
    if coeff.ndim == 3:
        v = lib().mo_interp3(coeff, coeff.shape[2], coeff.shape[1], coeff.shape[0], degree, p,
                             vmin, vmax, g)
    else:
        v = lib().mo_interp2(coeff, coeff.shape[1], coeff.shape[0], degree, p, vmin, vmax, g)
        g = g[:2]
    return (v, g) if with_grad else v


def spline_csize(n, rate, dim):
    return lib().mo_spline_csize(int(n), float(rate), int(dim))


def spline_scale(csize, n):
    return lib().mo_spline_scale(int(csize), int(n))


def spline_axis_table(n, csize):
    start = np.zeros(n, dtype=np.int32)
    nw = np.zeros(n, dtype=np.int32)
    w = np.zeros((n, 4))
    lib().mo_spline_axis_table(n, csize, start, w, nw)
    return start, w, nw


def spline_derivative_rows(n, csize):
    M = np.zeros((csize, n))
    lib().mo_spline_derivative_rows(n, csize, M)
    return M


def displacement(shape, csize, params):
    """Displacement field u (shape + (dim,)) for coefficient sizes csize (x, y[, z])."""
    dim = len(shape)
    s = _size3(shape)
    cs = np.array(list(csize) + [1] * (3 - len(csize)), dtype=np.int32)
    u = np.zeros(tuple(shape) + (dim,))
    lib().mo_displacement(dim, s, cs, np.ascontiguousarray(params, dtype=np.float64), u)
    return u


def displacement_at(shape, csize, params, pos):
    dim = len(shape)
    cs = np.array(list(csize) + [1] * (3 - len(csize)), dtype=np.int32)
    out = np.zeros(3)
    lib().mo_displacement_at(dim, _size3(shape), cs, np.ascontiguousarray(params, dtype=np.float64),
                             np.ascontiguousarray(pos, dtype=np.float64), out)
    return out[:dim]


def warp(moving, u, degree=3, with_grad=False):
    dim = moving.ndim
    coeff, vmin, vmax = prefilter(moving, degree)
    s = _size3(moving.shape)
    W = np.zeros(moving.shape)
    g = np.zeros(moving.shape + (dim,)) if with_grad else None
    lib().mo_warp(dim, s, coeff, degree, vmin, vmax, np.ascontiguousarray(u), W,
                  g.ctypes.data if with_grad else None)
    return (W, g) if with_grad else W


def image_gradient(W):
    dim = W.ndim
    g = np.zeros(W.shape + (dim,))
    lib().mo_image_gradient(dim, _size3(W.shape), np.ascontiguousarray(W, dtype=np.float64), g)
    return g


def ssd_value(W, R, norm=False, autothresh=0.0):
    return lib().mo_ssd_value(W.size, np.ascontiguousarray(W, dtype=np.float64),
                              np.ascontiguousarray(R, dtype=np.float32), int(norm), float(autothresh))


def ssd_force(W, R, grad, norm=False, autothresh=0.0):
    dim = W.ndim
    f = np.zeros(W.shape + (dim,))
    c = lib().mo_ssd_force(dim, W.size, np.ascontiguousarray(W, dtype=np.float64),
                           np.ascontiguousarray(R, dtype=np.float32),
                           np.ascontiguousarray(grad, dtype=np.float64), int(norm), float(autothresh), f)
    return c, f


def translate(force, csize, mode=0):
    """Adjoint: -B^T force.  force shape = image shape + (dim,)."""
    dim = force.shape[-1]
    shape = force.shape[:-1]
    cs = np.array(list(csize) + [1] * (3 - len(csize)), dtype=np.int32)
    p = np.zeros(dim * int(np.prod(csize)))
    lib().mo_translate(dim, _size3(shape), cs, np.ascontiguousarray(force, dtype=np.float64), mode, p)
    return p


def cost_grad(moving, fixed, csize, params, degree=3, norm=False, autothresh=0.0, grad_mode=0,
              weight=1.0, want_warped=False, want_force=False):
    dim = moving.ndim
    s = _size3(moving.shape)
    cs = np.array(list(csize) + [1] * (3 - len(csize)), dtype=np.int32)
    grad = np.zeros(dim * int(np.prod(csize)))
    W = np.zeros(moving.shape) if want_warped else None
    F = np.zeros(moving.shape + (dim,)) if want_force else None
    c = lib().mo_cost_grad(dim, s, np.ascontiguousarray(moving, dtype=np.float32),
                           np.ascontiguousarray(fixed, dtype=np.float32), cs,
                           np.ascontiguousarray(params, dtype=np.float64), degree, int(norm),
                           float(autothresh), grad_mode, float(weight), grad.ctypes.data,
                           W.ctypes.data if W is not None else None,
                           F.ctypes.data if F is not None else None)
    out = [c, grad]
    if want_warped:
        out.append(W)
    if want_force:
        out.append(F)
    return tuple(out)


def cost(moving, fixed, csize, params, degree=3, norm=False, autothresh=0.0):
    dim = moving.ndim
    cs = np.array(list(csize) + [1] * (3 - len(csize)), dtype=np.int32)
    return lib().mo_cost(dim, _size3(moving.shape), np.ascontiguousarray(moving, dtype=np.float32),
                         np.ascontiguousarray(fixed, dtype=np.float32), cs,
                         np.ascontiguousarray(params, dtype=np.float64), degree, int(norm),
                         float(autothresh))


class PrefilteredPair:
    """A moving image prefiltered once (FP64 coefficients and clamp range)
    with its fixed image, for many support-restricted cost evaluations."""

    def __init__(self, moving, fixed, csize, degree=3):
        self.dim = moving.ndim
        self.shape = moving.shape
        self.size = _size3(moving.shape)
        self.coeff, self.vmin, self.vmax = prefilter(moving, degree)
        self.fixed = np.ascontiguousarray(fixed, dtype=np.float32)
        self.csize = np.array(list(csize) + [1] * (3 - len(csize)), dtype=np.int32)
        self.degree = degree
        self.tables = [spline_axis_table(int(self.size[a]), int(self.csize[a]))[::2] for a in range(self.dim)]

    def support_box(self, k):
        """Voxel box [lo, hi) (x, y, z) of the support of coefficient k (its
        index in the coefficient grid, x fastest)."""
        lo = np.zeros(3, dtype=np.int32)
        hi = np.ones(3, dtype=np.int32)
        kk = [k % self.csize[0], (k // self.csize[0]) % self.csize[1], k // (self.csize[0] * self.csize[1])]
        for a in range(self.dim):
            start, nw = self.tables[a]
            inside = np.nonzero((start <= kk[a]) & (start + nw - 1 >= kk[a]))[0]
            if inside.size == 0:
                return None
            lo[a], hi[a] = inside[0], inside[-1] + 1
        return lo, hi

    def cost_box(self, params, lo, hi):
        return lib().mo_cost_box(self.dim, self.size, self.coeff, self.vmin, self.vmax, self.fixed, self.csize,
                                 np.ascontiguousarray(params, dtype=np.float64), self.degree,
                                 np.ascontiguousarray(lo, dtype=np.int32), np.ascontiguousarray(hi, dtype=np.int32))

    def fd(self, params, e, h):
        """Central difference (cost(p + h e_e) - cost(p - h e_e)) / 2h of the
        unnormalised cost in parameter e, over the support of its coefficient."""
        box = self.support_box(e // self.dim)
        if box is None:
            return 0.0
        p = np.array(params, dtype=np.float64)
        p[e] += h
        cp = self.cost_box(p, *box)
        p[e] -= 2 * h
        cm = self.cost_box(p, *box)
        return (cp - cm) / (2 * h)


def gauss_line(values, half_width):
    """Binomial filter of half width F with the reference's boundary taps."""
    x = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros_like(x)
    lib().mo_gauss_line(x, x.size, int(half_width), out)
    return out


def scale_line(values, n_out):
    x = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros(int(n_out))
    lib().mo_scale_line(x, x.size, out, int(n_out))
    return out


def scale_image(img, out_shape):
    """Resample a float image (numpy (z,y,x) or (y,x)) to out_shape."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros(tuple(out_shape), dtype=np.float32)
    lib().mo_scale_image(img, _size3(img.shape), _size3(out.shape), out)
    return out


def bspline3_deg(x, deg):
    return lib().mo_bspline3_deg(float(x), int(deg))


def integrate2(s1, s2, d1, d2, x0=0.0, L=8.0):
    return lib().mo_integrate2(float(s1), float(s2), int(d1), int(d2), float(x0), float(L))


def pp_cells(csize, rng, wdiv, wrot):
    """9^3 x 6 divcurl matrix cells (vxx, vxz, vxy, vyy, vyz, vzz); csize/rng (x,y,z)."""
    cells = np.zeros(9 * 9 * 9 * 6)
    lib().mo_pp_cells(np.asarray(csize, dtype=np.int32), np.asarray(rng, dtype=np.float32), float(wdiv),
                      float(wrot), cells)
    return cells


def divcurl(csize, cells, coef, with_grad=True):
    """Value (and gradient, 3 per coefficient) of the divcurl form; coef float32 (n, 3) x fastest."""
    c = np.ascontiguousarray(coef, dtype=np.float32).reshape(-1)
    g = np.zeros(c.size) if with_grad else None
    v = lib().mo_divcurl(np.asarray(csize, dtype=np.int32), np.ascontiguousarray(cells), c,
                         g.ctypes.data if with_grad else None)
    return (v, g) if with_grad else v
