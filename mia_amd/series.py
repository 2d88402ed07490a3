"""Batched series registration (libmia_host.so, include/mia_host.h).

The MI355X counterpart of `mia-2dmany2one-nonrigid`: many independent
frame-to-reference registrations (spline transformation, ssd cost), stepped in
lockstep through one batched GPU evaluation per iteration.
"""
import ctypes
import os
import re

import numpy as np

from ._lib import MiaHipError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmia_host.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "mia_host.h")

BATCHED_GRAPH = 0
BATCHED = 1
PER_FRAME = 2
DEVICE_BATCHED = 3
DEVICE_PER_FRAME = 4
DEVICE_SPLIT = 5

_vp = ctypes.c_void_p
SIGNATURES = {
    "mia_host_last_error": (ctypes.c_char_p, []),
    "mia_host_series_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_uint32), ctypes.c_float, ctypes.c_int,
                                              ctypes.POINTER(_vp)]),
    "mia_host_series_set_frame": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                                 ctypes.POINTER(ctypes.c_float)]),
    "mia_host_series_n_params": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                                ctypes.POINTER(ctypes.c_size_t)]),
    "mia_host_series_run": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, _vp, _vp, _vp,
                                           ctypes.POINTER(ctypes.c_double)]),
    "mia_host_series_destroy": (ctypes.c_int, [_vp]),
    "mia_host_nonrigid_register": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                                  ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                                  _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                                  ctypes.POINTER(ctypes.c_uint32), _vp, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_double)]),
    "mia_host_transform_save": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.c_char_p, _vp, ctypes.c_size_t, ctypes.c_char_p]),
    "mia_host_transform_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.POINTER(ctypes.c_uint32), _vp, ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p, ctypes.c_int,
                                               ctypes.c_char_p, ctypes.c_int]),
}


class LevelStats(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32 * 3), ("coeff_size", ctypes.c_uint32 * 3), ("evaluations", ctypes.c_int),
                ("cost", ctypes.c_double), ("seconds", ctypes.c_double), ("minimize_seconds", ctypes.c_double),
                ("loop_seconds", ctypes.c_double)]

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmia_host.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status == 0:
        return
    msg = load().mia_host_last_error().decode(errors="replace")
    if status == -1:
        raise ValueError(msg)
    if status == -2:
        raise RuntimeError(msg)
    raise MiaHipError(status, msg)


class SeriesRegistration:
    """frames: list of (src, ref) float32 numpy pairs of equal shape per pair."""

    def __init__(self, frames, rate=8.0, degree=3, device=0):
        self._L = load()
        self.dim = frames[0][0].ndim
        self.shapes = [f[0].shape for f in frames]
        n = len(frames)
        sz = np.zeros(3 * n, dtype=np.uint32)
        for i, s in enumerate(self.shapes):
            sz[3 * i:3 * i + self.dim] = s[::-1]
            if self.dim == 2:
                sz[3 * i + 2] = 1
        self._h = _vp()
        _check(self._L.mia_host_series_create(device, self.dim, n, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              float(rate), degree, ctypes.byref(self._h)))
        self.n = n
        self.slices = []
        for i, (src, ref) in enumerate(frames):
            src = np.ascontiguousarray(src, dtype=np.float32)
            ref = np.ascontiguousarray(ref, dtype=np.float32)
            if src.shape != ref.shape:
                raise RuntimeError("frame %d: src and ref differ in size" % i)
            _check(self._L.mia_host_series_set_frame(self._h, i, src.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                     ref.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
            a, b = ctypes.c_size_t(), ctypes.c_size_t()
            _check(self._L.mia_host_series_n_params(self._h, i, ctypes.byref(a), ctypes.byref(b)))
            self.slices.append(slice(b.value, b.value + a.value))
        self.n_params = self.slices[-1].stop

    def run(self, minimizer="lbfgs:maxiter=100", mode=BATCHED_GRAPH):
        """Returns (params, costs, evaluations, seconds)."""
        params = np.zeros(self.n_params)
        costs = np.zeros(self.n)
        evals = np.zeros(self.n, dtype=np.int32)
        sec = ctypes.c_double()
        _check(self._L.mia_host_series_run(self._h, minimizer.encode(), int(mode), params.ctypes.data,
                                           costs.ctypes.data, evals.ctypes.data, ctypes.byref(sec)))
        return params, costs, evals, sec.value

    def close(self):
        if self._h:
            self._L.mia_host_series_destroy(self._h)
            self._h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def normalize_pair(src, ref):
    """The joint mean-0 / sigma-1 map applied to each frame pair
    (mia/template/nonrigidregister.cxx:147-184), for callers that evaluate
    costs on the same images."""
    a = src.astype(np.float64)
    b = ref.astype(np.float64)
    n = 2 * a.size
    s = a.sum() + b.sum()
    s2 = (a * a).sum() + (b * b).sum()
    mean = s / n
    sigma = np.sqrt((s2 - s * s / n) / (n - 1))
    if sigma == 0.0:
        return src.astype(np.float32), ref.astype(np.float32)
    sa, sb = 1.0 / sigma, -mean / sigma
    return (a * sa + sb).astype(np.float32), (b * sa + sb).astype(np.float32)


def nonrigid_register(src, ref, transform="spline:rate=8", cost="image:cost=[ssd]", minimizer="lbfgs:maxiter=100",
                      levels=3, device=0):
    """Multiresolution registration of src to ref with the C++ driver
    (mia/template/nonrigidregister.cxx:192-296).  Returns (params,
    coefficient grid (x,y,z), per-level stats list of dicts, final cost)."""
    L = load()
    src = np.ascontiguousarray(src, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    if src.shape != ref.shape or src.ndim not in (2, 3):
        raise ValueError("nonrigid_register: src and ref must be 2D or 3D images of one shape")
    dim = src.ndim
    sz = np.array(list(src.shape[::-1]) + [1] * (3 - dim), dtype=np.uint32)
    # parameter count bound from the smallest rate named in the descriptor
    # (the spline creator's default rate is 10; anisorate may lower it)
    rates = [float(v) for v in re.findall(r"rate=\[?([0-9.eE+-]+)", transform)]
    rates += [float(v) for v in re.findall(r"anisorate=\[([^\]]*)\]", transform) for v in v.split(",") if float(v) > 0]
    r = max(1.0, min(rates) if rates else 10.0)
    cap = dim * int(np.prod([int((int(s) - 1 + r) / r) + 3 for s in sz[:dim]]))
    params = np.zeros(cap)
    n = ctypes.c_size_t()
    cs = (ctypes.c_uint32 * 3)()
    stats = (LevelStats * max(levels, 1))()
    fc = ctypes.c_double()
    fp = ctypes.POINTER(ctypes.c_float)
    _check(L.mia_host_nonrigid_register(int(device), dim, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                        src.ctypes.data_as(fp), ref.ctypes.data_as(fp), transform.encode(),
                                        cost.encode(), minimizer.encode(), int(levels), params.ctypes.data, cap,
                                        ctypes.byref(n), cs, ctypes.cast(stats, _vp), int(levels), ctypes.byref(fc)))
    st = [{"size": tuple(s.size), "coeff_size": tuple(s.coeff_size), "evaluations": s.evaluations,
           "cost": s.cost, "seconds": s.seconds, "minimize_seconds": s.minimize_seconds,
           "loop_seconds": s.loop_seconds} for s in stats[:levels]]
    return params[:n.value].copy(), tuple(cs), st, fc.value


def _pack_attributes(attributes):
    """{key: value} -> b"key\\0value\\0...\\0" (keys non-empty, no NUL inside)."""
    out = b""
    for k, v in sorted((attributes or {}).items()):
        k, v = str(k).encode(), str(v).encode()
        if not k or b"\0" in k or b"\0" in v:
            raise ValueError("attribute keys must be non-empty and keys and values free of NUL")
        out += k + b"\0" + v + b"\0"
    return out + b"\0"


def _unpack_attributes(buf):
    parts = buf.split(b"\0")
    out = {}
    k = 0
    while k + 1 < len(parts) and parts[k]:
        out[parts[k].decode()] = parts[k + 1].decode()
        k += 2
    return out


def save_transform(path, creator, shape, params, attributes=None):
    """Write a spline transformation file (MIA's xml layout).  shape: numpy
    image shape; creator: e.g. "spline:rate=8"; attributes: {str: str}."""
    L = load()
    dim = len(shape)
    sz = np.array(list(shape[::-1]) + [1] * (3 - dim), dtype=np.uint32)
    p = np.ascontiguousarray(params, dtype=np.float64)
    _check(L.mia_host_transform_save(str(path).encode(), dim, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                     creator.encode(), p.ctypes.data, p.size, _pack_attributes(attributes)))


def load_transform(path, dim, with_attributes=False):
    """Read a spline transformation file; returns (creator, numpy shape,
    coefficient grid (x,y,z), params[, attributes {str: str}])."""
    L = load()
    sz = (ctypes.c_uint32 * 3)()
    cs = (ctypes.c_uint32 * 3)()
    n = ctypes.c_size_t()
    cr = ctypes.create_string_buffer(1024)
    _check(L.mia_host_transform_load(str(path).encode(), int(dim), sz, cs, None, 0, ctypes.byref(n), None, 0,
                                     None, 0))
    params = np.zeros(n.value)
    attr = ctypes.create_string_buffer(1 << 16)
    _check(L.mia_host_transform_load(str(path).encode(), int(dim), sz, cs, params.ctypes.data, params.size,
                                     ctypes.byref(n), cr, 1024, attr, len(attr)))
    shape = tuple(int(v) for v in list(sz)[:dim][::-1])
    out = (cr.value.decode(), shape, tuple(cs), params)
    return out + (_unpack_attributes(attr.raw),) if with_attributes else out
