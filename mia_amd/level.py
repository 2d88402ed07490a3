"""Per-level GPU work of the multiresolution registration (include/mia_hip.h):
image resampling with the `scale` filter's semantics and the 3D divcurl
spline penalty."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def _u32(v):
    a = np.ascontiguousarray(v, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def scale_image(img, out_shape, device=0):
    """Resample a float image (numpy (z,y,x) or (y,x)) to out_shape
    (mia/3d/filter/scale.cc, mia/core/scaler1d.cc)."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    dim = img.ndim
    if dim not in (2, 3) or len(out_shape) != dim:
        raise ValueError("scale_image: 2D or 3D image and a matching output shape expected")
    out = np.zeros(tuple(int(v) for v in out_shape), dtype=np.float32)
    a, ap = _u32(list(img.shape[::-1]) + [1] * (3 - dim))
    b, bp = _u32(list(out.shape[::-1]) + [1] * (3 - dim))
    fp = ctypes.POINTER(ctypes.c_float)
    check(_lib.load().mia_hip_scale_image(int(device), dim, ap, img.ctypes.data_as(fp), bp,
                                          out.ctypes.data_as(fp)))
    return out


class DeviceImage:
    """A float image (numpy (z,y,x) or (y,x)) uploaded to the device
    (mia_hip_image): resampled there and set into contexts without host
    copies."""

    def __init__(self, img=None, device=0, _handle=None, _shape=None):
        self._L = _lib.load()
        if _handle is not None:
            self._h, self.shape = _handle, tuple(_shape)
            return
        img = np.ascontiguousarray(img, dtype=np.float32)
        if img.ndim not in (2, 3):
            raise ValueError("DeviceImage: 2D or 3D image expected")
        a, ap = _u32(list(img.shape[::-1]) + [1] * (3 - img.ndim))
        h = ctypes.c_void_p()
        check(self._L.mia_hip_image_create(int(device), img.ndim, ap, img.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           ctypes.byref(h)))
        self._h, self.shape = h, img.shape

    def scale(self, out_shape):
        """A new device image resampled to out_shape (mia_hip_scale_image's values)."""
        if len(out_shape) != len(self.shape):
            raise ValueError("DeviceImage.scale: dimension mismatch")
        b, bp = _u32(list(out_shape[::-1]) + [1] * (3 - len(out_shape)))
        h = ctypes.c_void_p()
        check(self._L.mia_hip_image_scale(self._h, bp, ctypes.byref(h)))
        return DeviceImage(_handle=h, _shape=out_shape)

    def read(self):
        out = np.zeros(self.shape, dtype=np.float32)
        check(self._L.mia_hip_image_read(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self._L.mia_hip_image_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DivcurlPenalty:
    """The divcurl form c^T P c over a coefficient grid csize (x,y,z) for an
    image of extent rng (x,y,z) (mia/3d/ppmatrix.cc); evaluate() returns the
    value and its gradient (3 per coefficient, unweighted)."""

    def __init__(self, csize, rng, div=1.0, curl=1.0, device=0):
        self._L = _lib.load()
        self.csize = tuple(int(v) for v in csize)
        cs, csp = _u32(self.csize)
        r = np.ascontiguousarray(rng, dtype=np.float32)
        self._h = ctypes.c_void_p()
        check(self._L.mia_hip_divcurl_create(int(device), csp, r.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             float(div), float(curl), ctypes.byref(self._h)))

    def evaluate(self, coef, with_grad=True):
        c = np.ascontiguousarray(coef, dtype=np.float32).reshape(-1)
        n = self.csize[0] * self.csize[1] * self.csize[2]
        if c.size != 3 * n:
            raise ValueError("divcurl: expected %d coefficient values" % (3 * n))
        v = ctypes.c_double()
        g = np.zeros(3 * n) if with_grad else None
        check(self._L.mia_hip_divcurl_evaluate(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               ctypes.byref(v), g.ctypes.data if with_grad else None))
        return (v.value, g) if with_grad else v.value

    def close(self):
        if self._h:
            self._L.mia_hip_divcurl_destroy(self._h)
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
