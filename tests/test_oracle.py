"""Pins the FP64 oracle (oracle/mia_oracle.c) before anything is compared to it.

Each test names what pins it: a known-answer vector from MIA's own unit tests
(tests/golden/reference_kat.json, with that test's tolerance), a closed form, an
invariant, a dense brute-force solve or a finite-difference derivative.  None of
them re-types the oracle's own formula.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def close_pct(a, b, pct):
    """BOOST_CHECK_CLOSE semantics: |a-b| <= pct% of both |a| and |b|."""
    a, b = float(a), float(b)
    d = abs(a - b)
    return d <= pct / 100.0 * abs(a) and d <= pct / 100.0 * abs(b) or d == 0.0


# --------------------------------------------------------------------------
# a1: kernel weights
# --------------------------------------------------------------------------

def test_bspline3_known_answers(kat):
    k = kat["bspline3_weights"]
    x = k["x"]
    start = O.fastfloor(x) - 1
    assert [start + i for i in range(4)] == k["index"]
    w = O.bspline3_weights(x - O.fastfloor(x))
    dw = O.bspline3_dweights(x - O.fastfloor(x))
    for a, b in zip(w, k["weights"]):
        assert close_pct(a, b, k["tol_percent"])
    for a, b in zip(dw, k["dweights"]):
        assert close_pct(a, b, k["tol_percent"])


def test_bspline3_closed_form_values():
    # beta_3 textbook values and support
    assert O.bspline3_value(0.0) == pytest.approx(2.0 / 3.0, abs=1e-15)
    assert O.bspline3_value(1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert O.bspline3_value(-1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert O.bspline3_value(0.5) == pytest.approx(23.0 / 48.0, abs=1e-15)
    assert O.bspline3_value(1.5) == pytest.approx(1.0 / 48.0, abs=1e-15)
    for x in (2.0, -2.0, 2.5, -3.0):
        assert O.bspline3_value(x) == 0.0


def test_bspline3_weight_at_cases(kat):
    """Weights equal beta_3 (value) or its FD derivative at the offsets the
    reference test lists (test_bspline.cc:280-331)."""
    k = kat["bspline3_weight_at"]
    for case in k["cases"]:
        x = case["x"]
        f = x - O.fastfloor(x)
        if case["kind"] == "value":
            w = O.bspline3_weights(f)
            ref = [O.bspline3_value(a) for a in case["beta_args"]]
        else:
            w = O.bspline3_dweights(f)
            h = 1e-6
            ref = [(O.bspline3_value(a + h) - O.bspline3_value(a - h)) / (2 * h) for a in case["beta_args"]]
        for a, b in zip(w, ref):
            assert abs(a - b) <= 1e-3 * max(abs(b), 1e-12) + 1e-9
    for x in k["beta_outside"]:
        assert O.bspline3_value(x) == 0.0


@pytest.mark.parametrize("f", [0.0, 0.1, 0.25, 0.5, 0.77, 0.999])
def test_bspline3_invariants(f):
    w = O.bspline3_weights(f)
    dw = O.bspline3_dweights(f)
    assert sum(w) == pytest.approx(1.0, abs=1e-15)          # partition of unity
    assert sum(dw) == pytest.approx(0.0, abs=1e-15)         # derivative of a constant
    # linear reproduction: sum_k (k-1) w_k = f
    assert sum((k - 1) * w[k] for k in range(4)) == pytest.approx(f, abs=1e-14)
    # weights are beta_3 at the distances to the 4 knots
    for k in range(4):
        assert w[k] == pytest.approx(O.bspline3_value(f - (k - 1)), abs=1e-14)


def test_fastfloor():
    for v in (0.0, 0.5, 3.25, -0.5, -3.75, 17.999, 100.5):
        assert O.fastfloor(v) == math.floor(v)
    # rounding to a multiple of 2^-16 before flooring
    assert O.fastfloor(3.0 - 2.0 ** -18) == 3
    assert O.fastfloor(-(2.0 ** -18)) == 0
    assert O.fastfloor(3.0 - 2.0 ** -15) == 2


@pytest.mark.parametrize("width", [1, 2, 3, 7, 12])
def test_mirror_index_brute_force(width):
    # mirrored sequence without edge duplication, period 2w-2
    if width == 1:
        seq = [0]
    else:
        seq = list(range(width)) + list(range(width - 2, 0, -1))
    for i in range(-3 * width - 2, 3 * width + 2):
        assert O.mirror_index(i, width) == seq[i % len(seq)]


# --------------------------------------------------------------------------
# a2/a3: prefilter and interpolation
# --------------------------------------------------------------------------

def _dense_interp_matrix(n):
    A = np.zeros((n, n))
    w = O.bspline3_weights(0.0)
    for i in range(n):
        for k in range(4):
            A[i, O.mirror_index(i - 1 + k, n)] += w[k]
    return A


def test_prefilter_matches_dense_solve(kat):
    k = kat["prefilter_mirror"]
    n = k["length"]
    A = _dense_interp_matrix(n)
    for j in range(-1, n):
        f = np.full(n, k["base_value"])
        if j >= 0:
            f[j] = k["spike_value"]
        c = O.filter_line(f)
        sol = np.linalg.lstsq(A, f, rcond=None)[0]
        for a, b in zip(c, sol):
            assert close_pct(a, b, k["tol_percent"])
        # tighter than the reference's 0.01 %: FP64 should agree to ~1e-12
        np.testing.assert_allclose(c, sol, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("n", [2, 3, 5, 33])
def test_prefilter_random_lines(n):
    rng = np.random.default_rng(n)
    f = rng.normal(size=n)
    A = _dense_interp_matrix(n)
    np.testing.assert_allclose(A @ O.filter_line(f), f, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("shape", [(7, 9), (6, 5, 8)])
def test_interpolation_reproduces_grid(shape):
    """mia/3d/test_interpol.cc:75-122: interpolation at grid points returns data."""
    rng = np.random.default_rng(3)
    img = rng.uniform(-1, 2, size=shape).astype(np.float32)
    for degree in (1, 3):
        coeff, vmin, vmax = O.prefilter(img, degree)
        assert vmin == min(0.0, float(img.min())) and vmax == float(img.max())
        for idx in np.ndindex(*shape):
            pos = np.array(idx[::-1], dtype=np.float64)
            v = O.interp(coeff, pos, degree, vmin, vmax)
            assert v == pytest.approx(float(img[idx]), abs=1e-10)


def test_linear_interpolation_half_points():
    """Linear data is reproduced exactly between grid points (test_interpol.cc:60-70)."""
    z, y, x = np.meshgrid(np.arange(5), np.arange(6), np.arange(7), indexing="ij")
    img = (x + y + z + 1).astype(np.float32)
    coeff, vmin, vmax = O.prefilter(img, 1)
    for p in [(0.5, 1, 2), (3, 2.5, 1), (2, 3, 3.5), (1.25, 4.75, 0.5)]:
        assert O.interp(coeff, np.array(p), 1, vmin, vmax) == pytest.approx(sum(p) + 1, abs=1e-12)


def test_interpolation_gradient_matches_fd():
    rng = np.random.default_rng(5)
    img = rng.uniform(0.2, 1.0, size=(8, 9, 10)).astype(np.float32)
    for degree in (1, 3):
        coeff, _, _ = O.prefilter(img, degree)
        for _ in range(10):
            p = rng.uniform(1.3, 6.7, size=3)
            v, g = O.interp(coeff, p, degree, -1e30, 1e30, with_grad=True)
            h = 1e-6
            for a in range(3):
                e = np.zeros(3)
                e[a] = h
                fd = (O.interp(coeff, p + e, degree, -1e30, 1e30) - O.interp(coeff, p - e, degree, -1e30, 1e30)) / (2 * h)
                assert g[a] == pytest.approx(fd, rel=1e-6, abs=1e-7)


def test_interpolation_clamp():
    img = np.zeros((6, 6), dtype=np.float32)
    img[2:4, 2:4] = 1.0
    coeff, vmin, vmax = O.prefilter(img, 3)
    assert vmin == 0.0 and vmax == 1.0
    # the cubic interpolant overshoots below 0 next to the block: clamped, zero gradient
    v, g = O.interp(coeff, np.array([0.5, 2.5]), 3, vmin, vmax, with_grad=True)
    raw = O.interp(coeff, np.array([0.5, 2.5]), 3, -1e30, 1e30)
    assert raw < 0.0 and v == 0.0 and np.all(g == 0.0)


# --------------------------------------------------------------------------
# a4: spline transformation geometry and displacement
# --------------------------------------------------------------------------

def test_csize_formula_and_throw(kat):
    # mia/3d/transform/spline.cc:130-143: rate 8 on 256 -> 32 + 2
    assert O.spline_csize(256, 8.0, 3) == 34
    assert O.spline_csize(128, 16.0, 2) == 10
    assert O.spline_csize(192, 8.0, 3) == 26      # (191 + 8) / 8 = 24.875 -> 24 + 2
    assert O.spline_csize(160, 8.0, 3) == 22
    k = kat["csize_throw_2d"]
    for rx, ry in k["rates_throw"]:
        assert O.spline_csize(k["size"][0], rx, 2) == -1 or O.spline_csize(k["size"][1], ry, 2) == -1


def test_axis_table_partition_of_unity():
    for n, cs in [(256, 34), (128, 10), (10, 5), (97, 19)]:
        start, w, nw = O.spline_axis_table(n, cs)
        assert start.min() >= 0 and (start + nw - 1).max() <= cs - 1
        for i in range(n):
            assert w[i, :nw[i]].sum() == pytest.approx(1.0, abs=1e-6)
        # the dropped fourth weight at the last sample is (numerically) zero
        assert abs(w[n - 1, 3]) < 1e-5


def test_spline_analytic_2d(kat):
    """test_spline.cc:36-82,168-182,202-217: a prefiltered sampled field is
    reproduced by the displacement."""
    k = kat["spline_analytic_2d"]
    rx, ry = k["range"]
    cx, cy = rx + 2, ry + 2
    ivx = (rx - 1) / (rx - 1)
    ivy = (ry - 1) / (ry - 1)

    def fx(x, y):
        vx, vy = x - rx / 2.0, y - ry / 2.0
        return math.exp(-(6 * vx * vx + 5 * vy * vy) / (rx * ry))

    def fy(x, y):
        vx, vy = x - rx / 2.0, y - ry / 2.0
        return math.exp(-(4 * vx * vx + 8 * vy * vy) / (rx * ry))

    field = np.zeros((cy, cx, 2))
    for y in range(cy):
        for x in range(cx):
            sx, sy = ivx * (x - 1), ivy * (y - 1)
            field[y, x] = (fx(sx, sy), fy(sx, sy))
    for x in range(cx):          # y lines first, as the fixture does
        for d in range(2):
            field[:, x, d] = O.filter_line(field[:, x, d])
    for y in range(cy):
        for d in range(2):
            field[y, :, d] = O.filter_line(field[y, :, d])
    params = field.reshape(-1)
    u = O.displacement((ry, rx), (cx, cy), params)
    px, py = k["probe"]
    assert close_pct(u[py, px, 0], fx(px, py), k["tol_percent_probe"])
    assert close_pct(u[py, px, 1], fy(px, py), k["tol_percent_probe"])
    for y in range(ry):
        for x in range(rx):
            assert close_pct(u[y, x, 0], fx(x, y), k["tol_percent_iterator"])
            assert close_pct(u[y, x, 1], fy(x, y), k["tol_percent_iterator"])


def test_spline_analytic_3d(kat):
    """mia/3d/transform/test_spline.cc:46-77,210-230 at the probe point."""
    k = kat["spline_analytic_3d"]
    rx, ry, rz = k["range"]
    sx, sy, sz = k["coeff_core"]
    iv = [(rx - 1) / (sx - 1), (ry - 1) / (sy - 1), (rz - 1) / (sz - 1)]
    cx, cy, cz = sx + 2, sy + 2, sz + 2
    center = np.array([rx / 2.0, ry / 2.0, rz / 2.0])
    fscale = rx * ry * rz
    coefs = [(6, 5, 4), (3, 7, 2), (2, 4, 5)]

    def f(v):
        X = v - center
        return np.array([math.exp(-(a * X[0] ** 2 + b * X[1] ** 2 + c * X[2] ** 2) / fscale) for a, b, c in coefs])

    zz, yy, xx = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    S = np.stack([iv[0] * (xx - 1), iv[1] * (yy - 1), iv[2] * (zz - 1)], axis=-1) - center
    field = np.stack([np.exp(-(a * S[..., 0] ** 2 + b * S[..., 1] ** 2 + c * S[..., 2] ** 2) / fscale)
                      for a, b, c in coefs], axis=-1)
    for axis in (2, 1, 0):       # x, then y, then z lines (set_coefficients_and_prefilter)
        field = np.apply_along_axis(O.filter_line, axis, field)
    params = field.reshape(-1)
    p = np.array(k["probe"])
    u = O.displacement_at((rz, ry, rx), (cx, cy, cz), params, p)
    ref = f(p)
    for a in range(3):
        assert close_pct(u[a], ref[a], k["tol_percent"])


# --------------------------------------------------------------------------
# a5: warp known answers
# --------------------------------------------------------------------------

def test_shift_warp_2d(kat):
    k = kat["spline_shift_warp_2d"]
    nx, ny = k["size"]
    cx, cy = k["coeff_size"]
    src = np.array(k["src"], dtype=np.float32).reshape(ny, nx)
    exp = np.array(k["expected"], dtype=np.float64).reshape(ny, nx)
    params = np.tile(np.array(k["shift"]), cx * cy)
    u = O.displacement((ny, nx), (cx, cy), params)
    np.testing.assert_allclose(u[..., 0], 1.0, atol=1e-6)
    np.testing.assert_allclose(u[..., 1], 2.0, atol=1e-6)
    W = O.warp(src, u, 3)
    for a, b in zip(W.reshape(-1), exp.reshape(-1)):
        if b > 0.001 or a > 0.001:
            assert close_pct(a, b, k["tol_percent"])


def test_shift_warp_3d(kat):
    k = kat["spline_shift_warp_3d"]
    k2 = kat["spline_shift_warp_2d"]
    nx, ny, nz = k["size"]
    cx, cy, cz = k["coeff_size"]
    src = np.zeros((nz, ny, nx), dtype=np.float32)
    src[k["src_slice"]] = np.array(k2["src"], dtype=np.float32).reshape(ny, nx)
    exp = np.array(k2["expected"], dtype=np.float64).reshape(ny, nx)
    params = np.tile(np.array(k["shift"]), cx * cy * cz)
    u = O.displacement((nz, ny, nx), (cx, cy, cz), params)
    W = O.warp(src, u, 3)
    for a, b in zip(W[k["result_slice"]].reshape(-1), exp.reshape(-1)):
        if b > 0.001 or a > 0.001:
            assert close_pct(a, b, k["tol_percent"])


# --------------------------------------------------------------------------
# a6/a7: SSD value, image gradient and force
# --------------------------------------------------------------------------

def test_ssd_2d_known_answer(kat):
    k = kat["ssd_2d"]
    nx, ny = k["size"]
    src = np.array(k["src"], dtype=np.float64).reshape(ny, nx)
    ref = np.array(k["ref"], dtype=np.float32).reshape(ny, nx)
    g = O.image_gradient(src)
    c, f = O.ssd_force(src, ref, g, norm=k["norm"])
    assert close_pct(O.ssd_value(src, ref, norm=k["norm"]), k["value"], k["tol_percent"])
    assert close_pct(c, k["value"], k["tol_percent"])
    x, y = k["force_at"]
    assert close_pct(f[y, x, 0], k["force"][0], k["tol_percent"])
    assert close_pct(f[y, x, 1], k["force"][1], k["tol_percent"])


def test_ssd_3d_known_answer(kat):
    k = kat["ssd_3d"]
    nx, ny, nz = k["size"]
    src = np.array(k["src"], dtype=np.float64).reshape(nz, ny, nx)
    ref = np.array(k["ref"], dtype=np.float32).reshape(nz, ny, nx)
    g = O.image_gradient(src)
    x, y, z = k["force_at"]
    assert close_pct(O.ssd_value(src, ref, norm=True), k["value_norm"], k["tol_percent"])
    assert close_pct(O.ssd_value(src, ref, norm=False), k["value"], k["tol_percent"])
    _, f = O.ssd_force(src, ref, g, norm=False)
    np.testing.assert_allclose(f[z, y, x], k["force"], rtol=1e-12)
    _, f = O.ssd_force(src, ref, g, norm=True)
    np.testing.assert_allclose(f[z, y, x], k["force_norm"], rtol=1e-12)


def test_image_gradient_border_rule():
    rng = np.random.default_rng(1)
    W = rng.normal(size=(4, 5, 6))
    g = O.image_gradient(W)
    assert np.all(g[:, :, 0, 0] == 0) and np.all(g[:, :, -1, 0] == 0)
    assert np.all(g[:, 0, :, 1] == 0) and np.all(g[:, -1, :, 1] == 0)
    assert np.all(g[0, :, :, 2] == 0) and np.all(g[-1, :, :, 2] == 0)
    np.testing.assert_allclose(g[1:-1, 1:-1, 1:-1, 0], 0.5 * (W[1:-1, 1:-1, 2:] - W[1:-1, 1:-1, :-2]))
    np.testing.assert_allclose(g[1:-1, 1:-1, 1:-1, 2], 0.5 * (W[2:, 1:-1, 1:-1] - W[:-2, 1:-1, 1:-1]))


def test_ssd_autothresh_mask():
    W = np.array([[0.1, 0.5], [0.9, 0.2]])
    R = np.zeros((2, 2), dtype=np.float32)
    # only 0.5 and 0.9 exceed 0.3: 0.5 * (0.25 + 0.81) / 2
    assert O.ssd_value(W, R, autothresh=0.3) == pytest.approx(0.5 * (0.25 + 0.81) / 2)
    assert O.ssd_value(W, R, autothresh=5.0) == pytest.approx(np.finfo(np.float32).max)


# --------------------------------------------------------------------------
# a8: translate (adjoint)
# --------------------------------------------------------------------------

def test_translate_constant_2d(kat):
    k = kat["translate_constant_2d"]
    rx, ry = 33, 65
    cx, cy = rx + 2, ry + 2
    force = np.tile(np.array(k["force"]), (ry, rx, 1))
    for mode in (0, 1):
        p = O.translate(force, (cx, cy), mode).reshape(cy, cx, 2)
        scale = ((rx - 1) / (cx - 3)) * ((ry - 1) / (cy - 3))
        for y in range(4, cy - 3):
            for x in range(4, cx - 3):
                assert close_pct(p[y, x, 0], -1.0 * scale, k["tol_percent"])
                assert close_pct(p[y, x, 1], -2.0 * scale, k["tol_percent"])


def test_translate_constant_3d(kat):
    k = kat["translate_constant_3d"]
    rx, ry, rz = 50, 80, 30
    cx, cy, cz = 42, 67, 27
    s = ((cx - 3) / (rx - 1)) * ((cy - 3) / (ry - 1)) * ((cz - 3) / (rz - 1))
    force = np.tile(np.array(k["force"]), (rz, ry, rx, 1))
    p = O.translate(force, (cx, cy, cz), 0).reshape(cz, cy, cx, 3)
    inner = p[3:cz - 2, 3:cy - 2, 3:cx - 2] * s
    for d in range(3):
        np.testing.assert_allclose(inner[..., d], -k["force"][d], rtol=k["tol_percent"] / 100)


def test_translate_is_adjoint_of_displacement():
    """<translate(f), dc> = -<f, B dc>: pins mode 1 as the exact transpose."""
    rng = np.random.default_rng(11)
    for shape, cs in [((23, 31), (7, 9)), ((13, 11, 17), (6, 5, 7))]:
        dim = len(shape)
        f = rng.normal(size=shape + (dim,))
        dc = rng.normal(size=dim * int(np.prod(cs)))
        lhs = O.translate(f, cs, 1) @ dc
        rhs = -(f * O.displacement(shape, cs, dc)).sum()
        assert lhs == pytest.approx(rhs, rel=1e-12)
        # the reference adjoint rows agree with the transpose up to float positions
        p0, p1 = O.translate(f, cs, 0), O.translate(f, cs, 1)
        assert np.abs(p0 - p1).max() <= 1e-5 * np.abs(p1).max()


def test_translate_fd_mock_cost_2d(kat):
    """test_spline.cc:697-725 with Cost2DMock: translate of the mock gradient
    equals central finite differences of the mock cost."""
    k = kat["translate_fd_2d"]
    nx, ny = k["size"]
    cx, cy = O.spline_csize(nx, k["rate"], 2), O.spline_csize(ny, k["rate"], 2)
    cxy = np.array([0.5 * nx, 0.5 * ny])
    r = math.sqrt(cxy[0] ** 2 + cxy[1] ** 2)
    yy, xx = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")

    def src(px, py):
        return np.exp(-((px - cxy[0]) ** 2 + (py - cxy[1]) ** 2) / r)

    ref = np.exp(-((xx - cxy[0] - 2) ** 2 + (yy - cxy[1] - 2) ** 2) / r)

    def cost(params):
        u = O.displacement((ny, nx), (cx, cy), params)
        v = src(xx - u[..., 0], yy - u[..., 1]) - ref
        return (v * v).sum()

    v = src(xx, yy) - ref
    sg = np.stack([-2.0 / r * (xx - cxy[0]) * src(xx, yy), -2.0 / r * (yy - cxy[1]) * src(xx, yy)], axis=-1)
    gradient = 2.0 * v[..., None] * sg
    params = np.zeros(2 * cx * cy)
    tr = O.translate(gradient, (cx, cy), 0)
    rng = np.random.default_rng(2)
    for i in rng.choice(params.size, 60, replace=False):
        fd = []
        for delta in (k["delta"], 1e-4):
            pp = params.copy()
            pp[i] += delta
            cp = cost(pp)
            pp[i] -= 2 * delta
            cm = cost(pp)
            fd.append((cp - cm) / (2 * delta))
        if abs(tr[i]) > 1e-9 or abs(fd[0]) > 1e-9:
            assert close_pct(tr[i], fd[0], k["tol_percent"])
        # the small step pins it much tighter than the reference's 5 %
        assert tr[i] == pytest.approx(fd[1], rel=1e-4, abs=1e-7)


# --------------------------------------------------------------------------
# a9: full cost + gradient; exact mode against finite differences
# --------------------------------------------------------------------------

def _smooth_pair(shape, seed):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    a = gaussian_filter(rng.normal(size=shape), 1.5, mode="mirror")
    b = gaussian_filter(rng.normal(size=shape), 1.5, mode="mirror")
    a = (a - a.min()) / (a.max() - a.min())
    b = (b - b.min()) / (b.max() - b.min())
    return a.astype(np.float32), b.astype(np.float32)


@pytest.mark.parametrize("shape,rate,degree", [((20, 24), 4.0, 3), ((14, 12, 16), 4.0, 3), ((14, 12, 16), 4.0, 1)])
def test_exact_gradient_matches_fd(shape, rate, degree):
    dim = len(shape)
    mov, fix = _smooth_pair(shape, 7)
    cs = tuple(O.spline_csize(n, rate, dim) for n in shape[::-1])
    rng = np.random.default_rng(9)
    params = rng.normal(scale=0.7, size=dim * int(np.prod(cs)))
    c, g = O.cost_grad(mov, fix, cs, params, degree=degree, grad_mode=1)
    assert c == pytest.approx(O.cost(mov, fix, cs, params, degree=degree), rel=1e-14)
    h = 1e-5
    for i in rng.choice(params.size, 25, replace=False):
        pp = params.copy()
        pp[i] += h
        cp = O.cost(mov, fix, cs, pp, degree=degree)
        pp[i] -= 2 * h
        cm = O.cost(mov, fix, cs, pp, degree=degree)
        fd = (cp - cm) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=2e-6 * np.abs(g).max())


@pytest.mark.parametrize("shape,rate", [((23, 29), 4.0), ((13, 11, 15), 3.0)])
def test_support_box_cost_difference_equals_full(shape, rate):
    """The support-restricted cost (used for the full-size FD check) against
    the whole-image cost: over the whole image it is the same sum, and for
    every coefficient (edges included) the difference of the box sums equals
    the difference of the full costs, so no voxel the coefficient moves lies
    outside its box.  The same box one voxel short at its high x end fails
    that check for every coefficient (the last voxel of a support carries a
    nonzero weight), so the check would catch a box that is too small."""
    dim = len(shape)
    mov, fix = _smooth_pair(shape, 5)
    cs = tuple(O.spline_csize(n, rate, dim) for n in shape[::-1])
    rng = np.random.default_rng(3)
    params = rng.normal(scale=0.8, size=dim * int(np.prod(cs)))
    pair = O.PrefilteredPair(mov, fix, cs)
    full = np.array(list(shape[::-1]) + [1] * (3 - dim), dtype=np.int32)
    assert pair.cost_box(params, np.zeros(3, np.int32), full) == pytest.approx(
        O.cost(mov, fix, cs, params), rel=1e-13)
    h = 0.05
    shrunk = 0
    for k in range(int(np.prod(cs))):
        e = dim * k + int(rng.integers(dim))
        pp = params.copy()
        pp[e] += h
        cp = O.cost(mov, fix, cs, pp)
        pp[e] -= 2 * h
        cm = O.cost(mov, fix, cs, pp)
        fd_full = (cp - cm) / (2 * h)
        assert pair.fd(params, e, h) == pytest.approx(fd_full, rel=1e-9, abs=1e-12), k
        lo, hi = pair.support_box(k)
        hi = hi.copy()
        hi[0] -= 1
        pp = params.copy()
        pp[e] += h
        a = pair.cost_box(pp, lo, hi)
        pp[e] -= 2 * h
        b = pair.cost_box(pp, lo, hi)
        shrunk += abs((a - b) / (2 * h) - fd_full) > 1e-9 * max(1.0, abs(fd_full))
    assert shrunk == np.prod(cs)


def test_reference_mode_is_warped_image_gradient():
    """grad_mode 0 is translate(central-difference force); at zero displacement
    the warped image is the moving image itself, so it also equals the exact
    gradient up to the difference between central and analytic derivatives."""
    mov, fix = _smooth_pair((18, 22), 3)
    cs = (O.spline_csize(22, 4.0, 2), O.spline_csize(18, 4.0, 2))
    params = np.zeros(2 * cs[0] * cs[1])
    c0, g0, W, F = O.cost_grad(mov, fix, cs, params, grad_mode=0, want_warped=True, want_force=True)
    np.testing.assert_allclose(W, mov, atol=1e-6)
    gW = O.image_gradient(W)
    np.testing.assert_allclose(F, gW * (W - fix)[..., None], atol=1e-12)
    np.testing.assert_allclose(g0, O.translate(F, cs, 0), rtol=1e-12, atol=1e-14)
    c1, g1 = O.cost_grad(mov, fix, cs, params, grad_mode=1)
    assert c0 == c1
    assert np.abs(g0 - g1).max() < 0.2 * np.abs(g1).max()


def test_weight_scales_cost_and_gradient():
    mov, fix = _smooth_pair((12, 14), 4)
    cs = (O.spline_csize(14, 4.0, 2), O.spline_csize(12, 4.0, 2))
    params = np.random.default_rng(0).normal(size=2 * cs[0] * cs[1])
    c1, g1 = O.cost_grad(mov, fix, cs, params)
    c2, g2 = O.cost_grad(mov, fix, cs, params, weight=0.25)
    assert c2 == pytest.approx(0.25 * c1)
    np.testing.assert_allclose(g2, 0.25 * g1)
