"""Oracle checks of the per-level restatements: the `scale` filter
(C1DScalar with the binomial prefilter) and the divcurl penalty (integrate2,
the P cells, the banded form).  Pinned by the reference's own known answers
(mia/core/test_scaler1d.cc, mia/3d/test_ppmatrix.cc), by closed forms of the
B-spline integrals and by properties of a symmetric quadratic form."""
import numpy as np
import pytest

from oracle import oracle as O


def _scaler_data():
    # mia/core/test_scaler1d.cc:108-114: 256 samples of 200 cos(2 pi x / 255)
    return 200 * np.cos(2 * np.pi / 255.0 * np.arange(256))


@pytest.mark.parametrize("target", [500, 130])
def test_scaler_reference_kat(target):
    """test_bspline3_upscale / test_bspline3_downscale
    (mia/core/test_scaler1d.cc:66-76,116-133): every output within 0.1 % of
    200 cos(2 pi i / (target - 1)) where either is above 1e-4."""
    r = O.scale_line(_scaler_data(), target)
    fx = 200 * np.cos(2 * np.pi / (target - 1) * np.arange(target))
    m = (np.abs(fx) > 1e-4) | (np.abs(r) > 1e-4)
    assert np.all(np.abs(fx[m] - r[m]) <= 1e-3 * np.abs(fx[m]))


def test_scaler_identity_and_constant():
    d = _scaler_data()
    np.testing.assert_array_equal(O.scale_line(d, 256), d)
    for n in (40, 97, 300):
        np.testing.assert_allclose(O.scale_line(np.full(128, 3.25), n), 3.25, rtol=1e-12)


def test_scaler_upscale_interpolates_samples():
    """Upscaling by an integer factor k on (n-1)k+1 samples passes through the
    input samples (cubic B-spline interpolation after the prefilter)."""
    rng = np.random.default_rng(1)
    d = rng.normal(size=33)
    r = O.scale_line(d, 32 * 4 + 1)
    np.testing.assert_allclose(r[::4], d, atol=1e-11)


def test_gauss_binomial_and_boundaries():
    # kernel: a delta in the interior gives row 2F of Pascal's triangle / 4^F
    F = 2
    d = np.zeros(21)
    d[10] = 1.0
    np.testing.assert_allclose(O.gauss_line(d, F)[8:13], np.array([1, 4, 6, 4, 1]) / 16.0)
    # constants pass unchanged (normalised kernel, reflected taps at both ends)
    np.testing.assert_allclose(O.gauss_line(np.full(17, 2.5), 3), 2.5, rtol=1e-14)
    # left end: x[-1] = x[0], x[-2] = x[1] (gauss.cc:84-96)
    x = np.arange(10, dtype=float) ** 2
    k = np.array([1, 4, 6, 4, 1]) / 16.0
    ext = np.concatenate([[x[1], x[0]], x, [x[-1], x[-2]]])
    np.testing.assert_allclose(O.gauss_line(x, 2), np.convolve(ext, k, mode="valid"), rtol=1e-14)


def test_scale_image_separable():
    """The 2D/3D filter is the 1D scaler along x, then y, then z."""
    rng = np.random.default_rng(2)
    img = rng.normal(size=(7, 9, 11)).astype(np.float32)
    out = O.scale_image(img, (13, 4, 20))
    a = img.astype(np.float64)
    b = np.stack([[O.scale_line(a[z, y], 20) for y in range(9)] for z in range(7)])
    c = np.stack([np.stack([O.scale_line(b[z, :, x], 4) for x in range(20)], axis=1) for z in range(7)])
    d = np.stack([np.stack([O.scale_line(c[:, y, x], 13) for x in range(20)], axis=1) for y in range(4)], axis=1)
    np.testing.assert_allclose(out, d.astype(np.float32), rtol=1e-6, atol=1e-6)


def test_bspline_derivatives_closed_form():
    x = np.linspace(-2.5, 2.5, 101)
    h = 1e-5
    for deg in (0, 1):
        num = np.array([(O.bspline3_deg(v + h, deg) - O.bspline3_deg(v - h, deg)) / (2 * h) for v in x])
        ana = np.array([O.bspline3_deg(v, deg + 1) for v in x])
        ok = np.abs(np.abs(x) - np.round(np.abs(x))) > 1e-3  # away from the knots
        np.testing.assert_allclose(num[ok], ana[ok], atol=1e-6)
    assert sum(O.bspline3_deg(v, 0) for v in (-1, 0, 1)) == pytest.approx(1.0)


def test_integrate2_closed_forms():
    """The autocorrelation of the cubic B-spline is the degree-7 B-spline:
    integral of B3(x) B3(x-k) = 151/315, 397/1680, 1/42, 1/5040 for k = 0..3;
    of B3'^2 = 2/3, of B3''^2 = 8/3, of B3 B3'' = -2/3.  Composite Simpson
    with 8 intervals per unit on these piecewise polynomials is accurate to
    a few 1e-5 (the reference integrates the same way)."""
    I = lambda d, a, b: O.integrate2(4.0, 4.0 + d, a, b, 0.0, 8.0)  # noqa: E731
    for k, exact in enumerate((151 / 315, 397 / 1680, 1 / 42)):
        assert I(k, 0, 0) == pytest.approx(exact, rel=5e-6)
        assert I(-k, 0, 0) == pytest.approx(exact, rel=5e-6)
    assert I(3, 0, 0) == pytest.approx(1 / 5040, rel=2e-3)
    assert I(0, 1, 1) == pytest.approx(2 / 3, rel=5e-5)
    assert I(0, 2, 2) == pytest.approx(8 / 3, rel=5e-5)
    assert I(0, 0, 2) == pytest.approx(-2 / 3, rel=5e-5)
    assert I(4, 0, 0) == 0.0 and I(-4, 2, 2) == 0.0
    # antisymmetry of the mixed first derivative integral
    assert I(1, 0, 1) == pytest.approx(-I(-1, 0, 1), rel=1e-12)


def _ppmatrix_fixture(dsize, rng, kind):
    """TransformSplineFixtureFieldBase::init (mia/3d/test_ppmatrix.cc:32-60):
    a (2 dsize + 1)^3 field sampled at step rng / dsize, prefiltered with the
    cubic kernel (the interpolator's coefficients), range 2 rng per axis."""
    n = 2 * dsize + 1
    ax = (np.arange(n) - dsize) * (rng / dsize)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    e = np.exp(-X * X - Y * Y - Z * Z)
    if kind == "div":
        comps = [X * e, Y * e, Z * e]
    coef = np.stack([O.prefilter(c.astype(np.float32), 3)[0] for c in comps], axis=-1).astype(np.float32)
    return n, coef


def test_divcurl_reference_kat_nocurl():
    """test_nocurl_bspline3_7_4 (mia/3d/test_ppmatrix.cc:132-150): for
    u = (x, y, z) exp(-r^2) the div form is 105 pi^1.5 / (sqrt(2) 8) within
    0.8 %, the div+curl form too, the curl form is 0 (1 + form within 4 %)."""
    n, coef = _ppmatrix_fixture(9, 4.0, "div")
    test = 105.0 * np.pi ** 1.5 / (np.sqrt(2.0) * 8.0)
    for wd, wr, expect, tol in ((1, 0, test, 0.008), (1, 1, test, 0.008)):
        v = O.divcurl((n, n, n), O.pp_cells((n, n, n), (8.0, 8.0, 8.0), wd, wr), coef, with_grad=False)
        assert v == pytest.approx(expect, rel=tol)
    v = O.divcurl((n, n, n), O.pp_cells((n, n, n), (8.0, 8.0, 8.0), 0, 1), coef, with_grad=False)
    assert 1.0 + v == pytest.approx(1.0, rel=0.04)


def test_divcurl_form_properties():
    """A symmetric quadratic form: value(0) = 0, c.g = 2 value, value(a c) =
    a^2 value, gradient equal to central differences, and the cells are
    symmetric under (dx, dy, dz) -> (-dx, -dy, -dz)."""
    cs = (7, 6, 5)
    rng = np.random.default_rng(3)
    cells = O.pp_cells(cs, (30.0, 25.0, 20.0), 1.0, 2.5)
    c = rng.normal(size=(np.prod(cs), 3)).astype(np.float32)
    assert O.divcurl(cs, cells, np.zeros_like(c), with_grad=False) == 0.0
    v, g = O.divcurl(cs, cells, c)
    assert np.dot(c.reshape(-1).astype(np.float64), g) == pytest.approx(2 * v, rel=1e-10)
    assert O.divcurl(cs, cells, 2 * c, with_grad=False) == pytest.approx(4 * v, rel=1e-10)
    P = cells.reshape(9, 9, 9, 6)
    np.testing.assert_allclose(P, P[::-1, ::-1, ::-1], rtol=1e-12, atol=1e-15)
    flat = c.reshape(-1).astype(np.float64)
    for k in rng.choice(flat.size, 12, replace=False):
        h = 1e-2
        cp = flat.copy()
        cp[k] += h
        cm = flat.copy()
        cm[k] -= h
        fd = (O.divcurl(cs, cells, cp.astype(np.float32), with_grad=False) -
              O.divcurl(cs, cells, cm.astype(np.float32), with_grad=False)) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max())


def test_divcurl_cells_linear_in_weights():
    """P(wdiv, wcurl) = wdiv P(1, 0) + wcurl P(0, 1): the cells are linear in
    the two weights (ppmatrix.cc:200-255)."""
    cs, r = (9, 8, 7), (31.0, 27.0, 22.0)
    a, b = O.pp_cells(cs, r, 1.0, 0.0), O.pp_cells(cs, r, 0.0, 1.0)
    np.testing.assert_allclose(O.pp_cells(cs, r, 0.3, 2.0), 0.3 * a + 2.0 * b, rtol=1e-12, atol=1e-14)
