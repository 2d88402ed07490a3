"""MI355X-native B-spline registration cost + gradient (MIA's spline/ssd path).

The compute path is libmia_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/mia_hip.h); this package is its Python view.
"""
from ._lib import ERR_LIMIT, GRAD_EXACT, GRAD_REFERENCE, MiaHipError  # noqa: F401
from .context import HipContext, spline_csize  # noqa: F401
