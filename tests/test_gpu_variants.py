"""GPU tests of the 3D warp tile variants and their autotuner: every variant
against the FP64 oracle (both interpolation degrees, including tiles whose
tap box exceeds the LDS pitches), batches whose instances run different
variants against one context per instance (bitwise), and the autotuner's
table."""
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
    cs = tuple(O.spline_csize(n, rate, 3) for n in shape[::-1])
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
    import mia_amd
    with mia_amd.HipContext(0) as c:
        with pytest.raises(ValueError):
            c.set_warp_variant(_n_variants())
        with pytest.raises(ValueError):
            c.set_warp_variant(-2)


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
