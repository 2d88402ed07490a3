"""The 3D cost + gradient workload at its real size: 256^3, spline rate 8
(34^3 x 3 = 117,912 parameters), cubic image interpolation, against the FP64
oracle, plus z-cell stage-1 rows between 160 and 256 voxels.

Inputs: Gaussian-smoothed white noise (sigma 4 voxels, seeds 1 and 2) rescaled
to [0, 1]; coefficients N(0, 1.5) clipped to +-4 voxels (seed 3).

Bounds (FP32 kernels, FP64 SSD sum and cross-block gradient sums, against
FP64):
  cost      relative 1e-6
  gradient  max-norm relative 1e-5 (|g - g_ref|_inf / |g_ref|_inf)
  warped    max abs 1e-5 (images in [0, 1])
Finite differences: central, h = 1e-3, of the FP64 oracle cost restricted to
each coefficient's support (the only voxels the coefficient moves), on 512
seeded coefficients (seed 4); the exact-mode GPU gradient must agree to 1e-5
of the gradient's max-norm.  Reference: mia/3d/transform/spline.cc:731-791
(translate), mia/3d/fullcost/image.cc:67-78 (evaluate).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (256, 256, 256)
RATE = 8.0
COST_RTOL = 1e-6
GRAD_RTOL = 1e-5
WARP_ATOL = 1e-5


def _rel_inf(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def headline():
    from mia_amd import synth
    print("generating 256^3 inputs", flush=True)
    fix = synth.smoothed_noise(SHAPE, 4.0, 1)
    mov = synth.smoothed_noise(SHAPE, 4.0, 2)
    cs = tuple(O.spline_csize(n, RATE, 3) for n in SHAPE[::-1])
    assert cs == (34, 34, 34)
    p = synth.random_spline_params(cs, 1.5, 4.0, 3)
    return mov, fix, cs, p


def _gpu(mov, fix, cs, p, mode, with_warp=False, s1path=-1):
    import mia_amd
    shape = mov.shape
    with mia_amd.HipContext(0) as c:
        c.set_stage1_path(s1path)
        c.setup(3, [shape], [cs], 3)
        c.set_images(0, mov, fix)
        c.set_cost(grad_mode=mode)
        path = c.stage1_path(0)
        costs, g = c.cost_grad(p)
        W = c.warp(0, p) if with_warp else None
    return costs[0], g, W, path


@pytest.mark.parametrize("mode,s1path", [(0, 1), (0, 0), (1, -1)])
def test_headline_vs_oracle(headline, mode, s1path):
    mov, fix, cs, p = headline
    print("256^3 rate 8 cubic, gradient mode", mode, "stage-1 layout", s1path, flush=True)
    cost, g, W, path = _gpu(mov, fix, cs, p, mode, with_warp=True, s1path=s1path)
    # the reference gradient in either stage-1 layout (z cell: every lane of
    # the four waves live, x = 255 on lane 63 of wave 3); the exact gradient
    # always per slice
    assert path == (s1path if mode == 0 else 0)
    c_ref, g_ref, W_ref = O.cost_grad(mov, fix, cs, p, grad_mode=mode, want_warped=True)
    e_c = abs(cost - c_ref) / abs(c_ref)
    e_g = _rel_inf(g, g_ref)
    e_w = np.abs(W - W_ref).max()
    print("cost %.12g oracle %.12g rel %.2e; gradient max-norm rel %.2e; warped max abs %.2e"
          % (cost, c_ref, e_c, e_g, e_w), flush=True)
    assert e_c < COST_RTOL
    assert e_g < GRAD_RTOL
    assert e_w < WARP_ATOL


def test_headline_fd_512(headline):
    mov, fix, cs, p = headline
    _, g, _, _ = _gpu(mov, fix, cs, p, 1)
    pair = O.PrefilteredPair(mov, fix, cs)
    idx = np.random.default_rng(4).choice(p.size, 512, replace=False)
    fd = np.array([pair.fd(p, int(e), 1e-3) for e in idx])
    gmax = np.abs(g).max()
    err = np.abs(g[idx] - fd)
    print("FD (h=1e-3, 512 coefficients): max |g - fd| / |g|_inf %.2e, median relative %.2e"
          % (err.max() / gmax, np.median(err / np.maximum(np.abs(fd), 1e-300))), flush=True)
    assert err.max() < GRAD_RTOL * gmax


@pytest.mark.parametrize("shape,rate,degree", [((36, 41, 256), 8.0, 3), ((30, 44, 200), 7.0, 3),
                                               ((28, 33, 256), 8.0, 1), ((40, 36, 161), 6.0, 3)])
def test_zcell_wide_rows(shape, rate, degree):
    """Rows of 161 to 256 voxels on the z-cell stage 1 (the widest it takes;
    y sizes chosen so that the y knot bands hold at most 9 rows, the z-cell
    register window)."""
    from mia_amd import synth
    import mia_amd
    print("case", shape, rate, degree, flush=True)
    fix = synth.smoothed_noise(shape, 3.0, 11)
    mov = synth.smoothed_noise(shape, 3.0, 12)
    cs = tuple(O.spline_csize(n, rate, 3) for n in shape[::-1])
    p = synth.random_spline_params(cs, 1.5, 4.0, 13)
    with mia_amd.HipContext(0) as c:
        c.set_stage1_path(1)
        c.setup(3, [shape], [cs], degree)
        c.set_images(0, mov, fix)
        assert c.stage1_path(0) == 1
        costs, g = c.cost_grad(p)
    c_ref, g_ref = O.cost_grad(mov, fix, cs, p, degree=degree)
    e_c = abs(costs[0] - c_ref) / abs(c_ref)
    e_g = _rel_inf(g, g_ref)
    print("cost rel %.2e, gradient max-norm rel %.2e" % (e_c, e_g), flush=True)
    assert e_c < COST_RTOL
    assert e_g < GRAD_RTOL
