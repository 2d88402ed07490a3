"""GPU tests of the tuned choices and their autotuner: every 3D and 2D warp
tile variant and both 3D stage-1 layouts against the FP64 oracle (both
interpolation degrees, including tiles whose tap box exceeds the LDS pitches
and images that the tiles do not divide), batches whose instances run
different choices against one context per instance (bitwise), and the
autotuner's table for each kind."""
import json

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

COST_RTOL = 1e-5
GRAD_RTOL = 1e-5
WARP_ATOL = 2e-5


def _inputs(shape, rate, seed, sd=1.5, clip=3.0):
    from mia_amd import synth
    fix = synth.smoothed_noise(shape, 2.0, seed)
    mov = synth.smoothed_noise(shape, 2.0, seed + 100)
    cs = tuple(O.spline_csize(n, rate, len(shape)) for n in shape[::-1])
    p = synth.random_spline_params(cs, sd, clip, seed + 200)
    return mov, fix, cs, p


def _n_variants():
    from mia_amd import autotune
    return len(autotune.variants())


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("shape,rate,degree,sd,clip", [((37, 44, 70), 6.0, 3, 1.5, 4.0),
                                                       ((29, 61, 33), 8.0, 1, 1.5, 3.0),
                                                       ((40, 40, 40), 5.0, 3, 4.0, 9.0)])
def test_variant_matches_oracle(variant, shape, rate, degree, sd, clip):
    import mia_amd
    mov, fix, cs, p = _inputs(shape, rate, 11, sd, clip)
    with mia_amd.HipContext(0) as c:
        c.set_warp_variant(variant)
        c.setup(3, [shape], [cs], degree)
        assert c.warp_variant(0) == variant
        c.set_images(0, mov, fix)
        costs, g = c.cost_grad(p)
        w = c.warp(0, p)
    c_ref, g_ref, w_ref = O.cost_grad(mov, fix, cs, p, degree=degree, want_warped=True)[:3]
    assert costs[0] == pytest.approx(c_ref, rel=COST_RTOL)
    assert np.abs(g - g_ref).max() <= GRAD_RTOL * np.abs(g_ref).max()
    assert np.abs(w - w_ref).max() <= WARP_ATOL


def test_variant_errors():
    """Out-of-range variants fail: beyond every dim's variants at once, beyond
    the 3D variants (but a 2D one) at the 3D set-up."""
    import mia_amd
    from mia_amd import autotune
    n2 = len(autotune.variants_2d())
    assert n2 > _n_variants()
    with mia_amd.HipContext(0) as c:
        with pytest.raises(ValueError):
            c.set_warp_variant(n2)
        with pytest.raises(ValueError):
            c.set_warp_variant(-2)
        c.set_warp_variant(_n_variants())
        with pytest.raises(ValueError):
            c.setup(3, [(20, 20, 20)], [(6, 6, 6)], 3)
        c.setup(2, [(40, 40)], [(8, 8)], 3)
        assert c.warp_variant(0) == _n_variants()


def test_autotune_and_mixed_variant_batch(tmp_path):
    """Autotune two buckets, then a table that maps four size buckets to four
    different variants: a batch with one instance per bucket runs four warp
    launches and matches one context per instance bitwise."""
    import mia_amd
    from mia_amd import _lib, autotune
    L = _lib.load()
    out = str(tmp_path / "tuned.json")
    try:
        table, us = autotune.run(out, degree=3, b_min=14, b_max=15)
        assert us.shape == (1, 2, _n_variants()) and np.all(us > 0)
        assert [r["bucket"] for r in table["table"]] == [14, 15]
        for r in table["table"]:
            assert r["variant"] == int(np.argmin(r["us"]))
        src, n = autotune.tuning_source()
        assert src == out and n == 2
        mixed = dict(table)
        mixed["table"] = [{"degree": 3, "bucket": b, "variant": v} for b, v in ((15, 1), (16, 2), (17, 3), (18, 0))]
        path = str(tmp_path / "mixed.json")
        with open(path, "w") as f:
            json.dump(mixed, f)
        _lib.check(L.mia_hip_load_warp_tuning(0, path.encode()))
        shapes = [(30, 33, 31), (40, 41, 39), (48, 50, 51), (60, 66, 63)]   # 2^15 .. 2^18 voxels
        data = [_inputs(s, 6.0, 30 + k, 1.5, 4.0) for k, s in enumerate(shapes)]
        with mia_amd.HipContext(0) as c:
            c.setup(3, shapes, [d[2] for d in data], 3)
            assert [c.warp_variant(i) for i in range(4)] == [1, 2, 3, 0]
            for i, d in enumerate(data):
                c.set_images(i, d[0], d[1])
            cb, gb = c.cost_grad(np.concatenate([d[3] for d in data]))
        off = 0
        for i, (s, d) in enumerate(zip(shapes, data)):
            with mia_amd.HipContext(0) as c1:
                c1.setup(3, [s], [d[2]], 3)
                c1.set_images(0, d[0], d[1])
                c1s, g1 = c1.cost_grad(d[3])
            assert c1s[0] == cb[i]
            assert np.array_equal(g1, gb[off:off + g1.size])
            off += g1.size
            c_ref, g_ref = O.cost_grad(d[0], d[1], d[2], d[3])
            assert c1s[0] == pytest.approx(c_ref, rel=COST_RTOL)
    finally:
        L.mia_hip_load_warp_tuning(0, None)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape,rate,degree", [((97, 131), 8.0, 3), ((50, 230), 6.0, 1), ((224, 96), 8.0, 3)])
def test_warp2d_variant_matches_oracle(variant, shape, rate, degree):
    import mia_amd
    mov, fix, cs, p = _inputs(shape, rate, 21, 1.5, 4.0)
    with mia_amd.HipContext(0) as c:
        c.set_warp_variant(variant)
        c.setup(2, [shape], [cs], degree)
        assert c.warp_variant(0) == variant
        c.set_images(0, mov, fix)
        costs, g = c.cost_grad(p)
        w = c.warp(0, p)
    with mia_amd.HipContext(0) as c0:
        c0.set_warp_variant(0)
        c0.setup(2, [shape], [cs], degree)
        c0.set_images(0, mov, fix)
        w0 = c0.warp(0, p)
    c_ref, g_ref, w_ref = O.cost_grad(mov, fix, cs, p, degree=degree, want_warped=True)[:3]
    print("variant", variant, "warped max diff to variant 0", float(np.abs(w - w0).max()))
    # every tile shape and tap source contracts in the same order: the same image
    assert np.array_equal(w, w0)
    assert costs[0] == pytest.approx(c_ref, rel=COST_RTOL)
    assert np.abs(g - g_ref).max() <= GRAD_RTOL * np.abs(g_ref).max()
    assert np.abs(w - w_ref).max() <= WARP_ATOL


@pytest.mark.parametrize("shape,rate,degree", [((44, 52, 61), 8.0, 3), ((44, 52, 61), 8.0, 1),
                                               ((32, 32, 32), 6.0, 3)])
def test_stage1_layouts_match_oracle(shape, rate, degree):
    """Both 3D stage-1 layouts of the reference gradient on instances that
    allow the z-cell layout: each against the oracle, and against each other
    to FP32 summation-order rounding.  32^3 alone in its context: three of
    the block's four waves lie beyond the row, and the image ends where the
    allocation does (edge-lane loads clamped to the row)."""
    import mia_amd
    mov, fix, cs, p = _inputs(shape, rate, 41, 1.5, 4.0)
    res = {}
    for path in (0, 1):
        with mia_amd.HipContext(0) as c:
            c.set_stage1_path(path)
            c.setup(3, [shape], [cs], degree)
            assert c.stage1_path(0) == path
            c.set_images(0, mov, fix)
            res[path] = c.cost_grad(p)
    c_ref, g_ref = O.cost_grad(mov, fix, cs, p, degree=degree)[:2]
    for path, (costs, g) in res.items():
        assert costs[0] == pytest.approx(c_ref, rel=COST_RTOL)
        assert np.abs(g - g_ref).max() <= GRAD_RTOL * np.abs(g_ref).max()
    gs, gz = res[0][1], res[1][1]
    print("layouts: grad max diff", float(np.abs(gs - gz).max()), "of", float(np.abs(g_ref).max()))
    assert np.abs(gs - gz).max() <= 1e-6 * np.abs(g_ref).max()


def test_stage1_path_errors():
    import mia_amd
    with mia_amd.HipContext(0) as c:
        for bad in (-2, 2):
            with pytest.raises(ValueError):
                c.set_stage1_path(bad)


def test_autotune_2d_and_stage1_tables(tmp_path):
    """Tune the 2D warp tile and the stage-1 layout into one table (each
    tuning keeps the other kind's entries), check the library reports the
    architecture it is keyed by, then load tables that map size buckets to
    different 2D variants and stage-1 layouts: batches with one instance per
    bucket match one context per instance bitwise."""
    import mia_amd
    from mia_amd import _lib, autotune
    L = _lib.load()
    out = str(tmp_path / "tuned.json")
    try:
        t2, us2 = autotune.run(out, degree=3, b_min=12, b_max=13, kind="warp2d")
        assert us2.shape == (1, 2, len(autotune.variants_2d())) and np.all(us2 > 0)
        t3, us3 = autotune.run(out, degree=3, b_min=15, b_max=15, kind="stage1_3d")
        assert us3.shape == (1, 1, 2) and np.all(us3 > 0)
        assert t3["arch"] == autotune.device_arch(0) and t3["arch"].startswith("gfx")
        kinds = [r["kernel"] for r in t3["table"]]
        assert kinds.count("warp2d") == 2 and kinds.count("stage1_3d") == 1
        for r in t3["table"]:
            assert r["variant"] == int(np.argmin(r["us"]))
        src, n = autotune.tuning_source()
        assert src == out and n == 3

        mixed = {"arch": t3["arch"], "table": [
            {"kernel": "warp2d", "degree": 3, "bucket": b, "variant": v} for b, v in ((13, 1), (14, 2), (15, 3), (16, 0))
        ] + [{"kernel": "stage1_3d", "degree": 3, "bucket": b, "variant": v} for b, v in ((15, 0), (16, 1))]}
        path = str(tmp_path / "mixed.json")
        with open(path, "w") as f:
            json.dump(mixed, f)
        _lib.check(L.mia_hip_load_warp_tuning(0, path.encode()))
        cases = {2: [(70, 101), (110, 130), (181, 170), (250, 251)],   # 2^13 .. 2^16 pixels
                 3: [(30, 33, 31), (40, 41, 39)]}                       # 2^15, 2^16 voxels
        for dim, shapes in cases.items():
            rate = 8.0
            data = [_inputs(s, rate, 60 + k, 1.5, 4.0) for k, s in enumerate(shapes)]
            with mia_amd.HipContext(0) as c:
                c.setup(dim, shapes, [d[2] for d in data], 3)
                if dim == 2:
                    assert [c.warp_variant(i) for i in range(4)] == [1, 2, 3, 0]
                else:
                    assert [c.stage1_path(i) for i in range(2)] == [0, 1]
                for i, d in enumerate(data):
                    c.set_images(i, d[0], d[1])
                cb, gb = c.cost_grad(np.concatenate([d[3] for d in data]))
            off = 0
            for i, (s, d) in enumerate(zip(shapes, data)):
                with mia_amd.HipContext(0) as c1:
                    c1.setup(dim, [s], [d[2]], 3)
                    c1.set_images(0, d[0], d[1])
                    c1s, g1 = c1.cost_grad(d[3])
                assert c1s[0] == cb[i]
                assert np.array_equal(g1, gb[off:off + g1.size])
                off += g1.size
                c_ref, g_ref = O.cost_grad(d[0], d[1], d[2], d[3])[:2]
                assert c1s[0] == pytest.approx(c_ref, rel=COST_RTOL)
                assert np.abs(g1 - g_ref).max() <= GRAD_RTOL * np.abs(g_ref).max()
    finally:
        L.mia_hip_load_warp_tuning(0, None)


def test_mixed_2d_batch_one_timed_variant(tmp_path):
    """With per-variant timings in the table, a 2D batch whose instances fall
    into buckets with different tuned variants runs one warp launch: the
    variant with the least timing summed over the instances' buckets (here
    variant 2, which no bucket picks by itself).  Every variant computes the
    same image, so the batch still matches one context per instance, each
    with its own bucket's variant, bitwise."""
    import mia_amd
    from mia_amd import _lib, autotune
    L = _lib.load()
    nv = len(autotune.variants_2d())
    try:
        arch = autotune.device_arch(0)
        rows = []
        for b, best in ((13, 1), (14, 3), (15, 0), (16, 4)):
            us = [10.0] * nv
            us[best] = 7.0
            us[2] = 8.0
            rows.append({"kernel": "warp2d", "degree": 3, "bucket": b, "variant": best, "us": us})
        path = str(tmp_path / "timed.json")
        with open(path, "w") as f:
            json.dump({"arch": arch, "table": rows}, f)
        _lib.check(L.mia_hip_load_warp_tuning(0, path.encode()))
        shapes = [(70, 101), (110, 130), (181, 170), (250, 251)]   # 2^13 .. 2^16 pixels
        data = [_inputs(s, 8.0, 70 + k, 1.5, 4.0) for k, s in enumerate(shapes)]
        with mia_amd.HipContext(0) as c:
            c.setup(2, shapes, [d[2] for d in data], 3)
            assert [c.warp_variant(i) for i in range(4)] == [2, 2, 2, 2]
            for i, d in enumerate(data):
                c.set_images(i, d[0], d[1])
            cb, gb = c.cost_grad(np.concatenate([d[3] for d in data]))
            wb = [c.warp(i, d[3]) for i, d in enumerate(data)]
        off = 0
        for i, (s, d) in enumerate(zip(shapes, data)):
            with mia_amd.HipContext(0) as c1:
                c1.setup(2, [s], [d[2]], 3)
                assert c1.warp_variant(0) == [1, 3, 0, 4][i]
                c1.set_images(0, d[0], d[1])
                c1s, g1 = c1.cost_grad(d[3])
                w1 = c1.warp(0, d[3])
            assert c1s[0] == cb[i]
            assert np.array_equal(g1, gb[off:off + g1.size])
            assert np.array_equal(w1, wb[i])
            off += g1.size
    finally:
        L.mia_hip_load_warp_tuning(0, None)
