"""GPU tests of the per-level work through the C ABI (image resampling with
the `scale` filter's semantics, the divcurl penalty) against the FP64
oracle, and the C++ host-layer GPU test program (full cost, cost list,
minimizer problem with penalty, warp, multiresolution registrations)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _img(shape, seed):
    from mia_amd import synth
    return synth.smoothed_noise(shape, 2.0, seed)


@pytest.mark.parametrize("a,b", [((80, 100), (40, 50)), ((45, 60), (150, 170)), ((64, 64), (17, 64)),
                                 ((32, 36, 40), (16, 18, 20)), ((16, 20, 24), (33, 20, 48)),
                                 ((40, 48, 48), (10, 12, 12))])
def test_scale_image_matches_oracle(a, b):
    from mia_amd import level
    img = _img(a, 3)
    out = level.scale_image(img, b)
    ref = O.scale_image(img, b)
    assert out.shape == tuple(b)
    assert np.abs(out - ref).max() <= 1e-6 * np.abs(ref).max()


def test_scale_kat_cosine_rows():
    """The reference's 1D scaler test (mia/core/test_scaler1d.cc:116-133) on
    rows of a 2D image: 256 -> 500 and 256 -> 130 within 0.1 %."""
    from mia_amd import level
    row = (200 * np.cos(2 * np.pi / 255.0 * np.arange(256))).astype(np.float32)
    img = np.stack([row, row, row])
    for tgt in (500, 130):
        out = level.scale_image(img, (3, tgt))
        fx = 200 * np.cos(2 * np.pi / (tgt - 1) * np.arange(tgt))
        for r in out:
            m = np.abs(fx) > 1e-2
            assert np.all(np.abs(r[m] - fx[m]) <= 1e-3 * np.abs(fx[m]) + 1e-4)


def test_scale_errors():
    from mia_amd import level
    with pytest.raises(RuntimeError):
        level.scale_image(np.ones((8, 8), np.float32), (8, 1))
    with pytest.raises(ValueError):
        level.scale_image(np.ones((8,), np.float32), (4,))


@pytest.mark.parametrize("cs,rng,w", [((13, 11, 9), (40.0, 36.0, 30.0), (1.0, 1.0)),
                                      ((10, 10, 10), (32.0, 32.0, 32.0), (1.0, 0.0)),
                                      ((26, 26, 22), (192.0, 192.0, 160.0), (0.3, 2.0))])
def test_divcurl_matches_oracle(cs, rng, w):
    from mia_amd import level
    r = np.random.default_rng(5)
    c = r.normal(size=(int(np.prod(cs)), 3)).astype(np.float32)
    v_ref, g_ref = O.divcurl(cs, O.pp_cells(cs, rng, *w), c)
    with level.DivcurlPenalty(cs, rng, *w) as p:
        v, g = p.evaluate(c)
        v2 = p.evaluate(c, with_grad=False)
    assert v == pytest.approx(v_ref, rel=1e-10)
    assert v2 == v
    assert np.abs(g - g_ref).max() <= 1e-11 * np.abs(g_ref).max()


def test_divcurl_reference_kat_on_gpu():
    """test_nocurl_bspline3_7_4 (mia/3d/test_ppmatrix.cc:132-150) on the GPU."""
    from mia_amd import level
    n, dsize, rngv = 19, 9, 4.0
    ax = (np.arange(n) - dsize) * (rngv / dsize)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    e = np.exp(-X * X - Y * Y - Z * Z)
    coef = np.stack([O.prefilter(cc.astype(np.float32), 3)[0] for cc in (X * e, Y * e, Z * e)], axis=-1)
    test = 105.0 * np.pi ** 1.5 / (np.sqrt(2.0) * 8.0)
    with level.DivcurlPenalty((n, n, n), (8.0, 8.0, 8.0), 1.0, 0.0) as p:
        assert p.evaluate(coef, with_grad=False) == pytest.approx(test, rel=0.008)


def test_host_layer_gpu_cpp():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_nonrigid_gpu")
    assert os.path.exists(exe), "tests/cpp/bin/test_nonrigid_gpu is not built: run __graft_entry__.build()"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr[-2000:]
    assert out.stdout.strip().endswith("OK")


def test_nonrigid_register_python_2d():
    """Three-level 2D registration through the C ABI of the C++ driver:
    per-level sizes 1/4, 1/2, 1 and a final cost well below the identity's."""
    from mia_amd import series, synth
    fix = synth.gaussian_blobs((96, 112), seed=7)
    import mia_amd
    cs = tuple(mia_amd.spline_csize(n, 16.0, 2) for n in fix.shape[::-1])
    p = synth.random_spline_params(cs, 1.5, 3.0, 8)
    mov = synth.spline_warp(fix, cs, p).astype(np.float32)
    params, grid, stats, final = series.nonrigid_register(mov, fix, "spline:rate=8", "image:cost=[ssd]",
                                                          "lbfgs:maxiter=60", levels=3)
    assert [s["size"][:2] for s in stats] == [(28, 24), (56, 48), (112, 96)]
    assert grid[:2] == (mia_amd.spline_csize(112, 8.0, 2), mia_amd.spline_csize(96, 8.0, 2))
    assert params.size == 2 * grid[0] * grid[1]
    a, b = series.normalize_pair(mov, fix)
    c_id = 0.5 * float(np.sum((a.astype(np.float64) - b) ** 2))
    assert final < 0.3 * c_id
    with pytest.raises(ValueError):
        series.nonrigid_register(mov, fix, "spline:rate=8", "image:cost=[ncc]", "lbfgs:maxiter=5", levels=1)


def test_penalty_in_context_equals_host_sum():
    """The divcurl penalty attached to a context: every evaluation (eager and
    the captured graph) returns the image cost plus weight * value and the
    image gradient plus -weight * the penalty's gradient, bitwise as the
    host sums them (TNonrigRegGradientProblem::evaluate_fdf,
    mia/template/nonrigidregister.cxx:422-429); detached, the image cost alone."""
    import mia_amd
    from mia_amd import level, synth
    shape = (34, 30, 38)
    fix = synth.smoothed_noise(shape, 2.0, 3)
    mov = synth.smoothed_noise(shape, 2.0, 4)
    cs = tuple(mia_amd.spline_csize(n, 6.0, 3) for n in shape[::-1])
    p = synth.random_spline_params(cs, 1.5, 3.0, 5)
    w = 0.0137
    with mia_amd.HipContext(0) as c, level.DivcurlPenalty(cs, tuple(float(v) for v in shape[::-1])) as pen:
        c.setup(3, [shape], [cs], 3)
        c.set_images(0, mov, fix)
        c_img, g_img = c.cost_grad(p)
        v, g = pen.evaluate(p.astype(np.float32))
        c.set_penalty(0, pen, w)
        c1, g1 = c.cost_grad(p)
        assert c1[0] == c_img[0] + w * v
        assert np.array_equal(g1, g_img + (-w * g))
        c.graph_capture()
        for _ in range(3):
            cg, gg = c.graph_run(p)
            assert cg[0] == c1[0] and np.array_equal(gg, g1)
        with pytest.raises(RuntimeError):
            with level.DivcurlPenalty((cs[0] + 1, cs[1], cs[2]), (1.0, 1.0, 1.0)) as bad:
                c.set_penalty(0, bad, w)
        c.set_penalty(0, None, 0.0)
        c2, g2 = c.cost_grad(p)
        assert c2[0] == c_img[0] and np.array_equal(g2, g_img)


def test_nonrigid_register_device_matches_host_3d():
    """Two-level 3D registration with the divcurl penalty: the on-device
    L-BFGS (evaluation, penalty and optimizer step in one captured graph per
    ten iterations) against the host L-BFGS driving the same evaluations
    ("host=1"); same algorithm, different summation order of its dot
    products, so the runs agree to rounding."""
    import mia_amd
    from mia_amd import series, synth
    shape = (40, 44, 48)
    fix = synth.smoothed_noise(shape, 3.0, 31)
    cs = tuple(mia_amd.spline_csize(n, 8.0, 3) for n in shape[::-1])
    q = synth.random_spline_params(cs, 1.5, 3.0, 32)
    mov = synth.spline_warp(fix, cs, q)
    tr = "spline:rate=8,penalty=[divcurl:weight=0.01]"
    pd, gd, sd, fd = series.nonrigid_register(mov, fix, tr, "image:cost=[ssd]", "lbfgs:maxiter=8", levels=2)
    ph, gh, sh, fh = series.nonrigid_register(mov, fix, tr, "image:cost=[ssd]", "lbfgs:maxiter=8,host=1", levels=2)
    print("device", [(s["evaluations"], s["cost"]) for s in sd], fd)
    print("host  ", [(s["evaluations"], s["cost"]) for s in sh], fh)
    assert gd == gh
    assert [s["evaluations"] for s in sd] == [s["evaluations"] for s in sh]
    for a, b in zip(sd, sh):
        assert a["cost"] == pytest.approx(b["cost"], rel=1e-9)
    assert fd == pytest.approx(fh, rel=1e-9)
    assert np.abs(pd - ph).max() <= 1e-7 * max(1.0, np.abs(ph).max())


def test_nonrigid_register_matches_cpu_port():
    """The device driver against the CPU port's multiresolution registration
    (the bench's CPU baseline: FP32 MIA data flow, the same L-BFGS, divcurl
    evaluated in FP64 on the host) on the same two-level problem: the same
    level sizes, coefficient grids and evaluation counts, and final costs equal
    to FP32 evaluation rounding."""
    import ctypes
    import mia_amd
    from mia_amd import series, synth
    shape = (40, 44, 48)
    fix = synth.smoothed_noise(shape, 3.0, 31)
    cs = tuple(mia_amd.spline_csize(n, 8.0, 3) for n in shape[::-1])
    q = synth.random_spline_params(cs, 1.5, 3.0, 32)
    mov = synth.spline_warp(fix, cs, q)
    tr = "spline:rate=8,penalty=[divcurl:weight=0.01]"
    pd, gd, sd, fd = series.nonrigid_register(mov, fix, tr, "image:cost=[ssd]", "lbfgs:maxiter=8", levels=2)
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "lib", "libmia_cpu_port.so"))
    fp = np.ctypeslib.ndpointer(np.float32, flags="C")
    L.mcp_register_3d.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.mcp_register_3d.restype = ctypes.c_double
    lsec = np.zeros(2)
    lev = np.zeros(2, dtype=np.int32)
    lcost = np.zeros(2)
    cf = ctypes.c_double()
    ccs = np.zeros(3, dtype=np.int32)
    L.mcp_register_3d(np.ascontiguousarray(mov, dtype=np.float32), np.ascontiguousarray(fix), shape[2], shape[1],
                      shape[0], 8.0, 2, 8, 0.01, 4, lsec.ctypes.data, lev.ctypes.data, lcost.ctypes.data, ctypes.byref(cf),
                      ccs.ctypes.data)
    print("device", [s["evaluations"] for s in sd], sd[-1]["cost"], "cpu port", list(lev), cf.value)
    assert tuple(int(v) for v in ccs) == tuple(gd)
    assert [s["evaluations"] for s in sd] == [int(v) for v in lev]
    for s, c in zip(sd, lcost):
        assert s["cost"] == pytest.approx(c, rel=1e-5)
    assert sd[-1]["cost"] == pytest.approx(cf.value, rel=1e-5)


def test_setup_grid_keeps_images():
    """mia_hip_setup_grid: new coefficient grids for the same images (as a
    refinement does) give bitwise the cost and gradient of a context set up
    afresh with those grids and the images set again; 3D and 2D, a batch of
    two instances; the grid must be set up first."""
    import mia_amd
    from mia_amd import synth
    for dim, shapes, rates in ((3, [(30, 36, 41), (25, 33, 28)], (9.0, 6.0)), (2, [(70, 90), (64, 50)], (12.0, 7.0))):
        data = [(synth.smoothed_noise(s, 2.0, 80 + k), synth.smoothed_noise(s, 2.0, 90 + k)) for k, s in enumerate(shapes)]
        cs0 = [tuple(mia_amd.spline_csize(n, rates[0], dim) for n in s[::-1]) for s in shapes]
        cs1 = [tuple(mia_amd.spline_csize(n, rates[1], dim) for n in s[::-1]) for s in shapes]
        p1 = np.concatenate([synth.random_spline_params(c, 1.0, 3.0, 95 + k) for k, c in enumerate(cs1)])
        with mia_amd.HipContext(0) as c:
            c.setup(dim, shapes, cs0, 3)
            for i, (m, f) in enumerate(data):
                c.set_images(i, m, f)
            c.cost_grad(np.zeros(c.n_params_total))
            c.setup_grid(cs1)
            assert c.n_params_total == p1.size
            ca, ga = c.cost_grad(p1)
        with mia_amd.HipContext(0) as c:
            c.setup(dim, shapes, cs1, 3)
            for i, (m, f) in enumerate(data):
                c.set_images(i, m, f)
            cb, gb = c.cost_grad(p1)
        assert np.array_equal(ca, cb) and np.array_equal(ga, gb), dim
    with mia_amd.HipContext(0) as c:
        with pytest.raises(RuntimeError):
            c.setup_grid([(8, 8, 8)])


def test_device_images_match_host_path():
    """Device-resident images (mia_hip_image): an upload reads back bitwise;
    resampling on the device gives bitwise mia_hip_scale_image's values
    (down and up, 3D and 2D); a context set from device images evaluates
    bitwise as one set from host images; size mismatches are errors."""
    import mia_amd
    from mia_amd import level, synth
    for shape, out in (((40, 48, 56), (20, 24, 28)), ((33, 30, 27), (40, 41, 29)), ((60, 70), (30, 35))):
        img = synth.smoothed_noise(shape, 2.0, 7).astype(np.float32)
        d = level.DeviceImage(img)
        assert np.array_equal(d.read(), img)
        ds = d.scale(out)
        assert np.array_equal(ds.read(), level.scale_image(img, out))
    shape = (30, 36, 41)
    mov, fix = synth.smoothed_noise(shape, 2.0, 11), synth.smoothed_noise(shape, 2.0, 12)
    cs = tuple(mia_amd.spline_csize(n, 8.0, 3) for n in shape[::-1])
    p = synth.random_spline_params(cs, 1.0, 3.0, 13)
    dm, df = level.DeviceImage(mov), level.DeviceImage(fix)
    with mia_amd.HipContext(0) as c:
        c.setup(3, [shape], [cs], 3)
        c.set_images_image(0, dm, df)
        ca, ga = c.cost_grad(p)
        with pytest.raises(RuntimeError):
            c.set_images_image(0, dm.scale((15, 18, 20)), df)
    with mia_amd.HipContext(0) as c:
        c.setup(3, [shape], [cs], 3)
        c.set_images(0, mov, fix)
        cb, gb = c.cost_grad(p)
    assert np.array_equal(ca, cb) and np.array_equal(ga, gb)
