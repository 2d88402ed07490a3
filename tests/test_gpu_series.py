"""GPU tests of the batched series registration (libmia_host over libmia_hip).

* batched (graph), batched (eager) and one-context-per-frame runs give
  bitwise identical parameters and final costs ("equal final SSD");
* MIA's gdsq minimizer driven by the GPU cost/gradient follows the same
  trajectory as gdsq driven by the FP64 oracle (restated below from
  mia/core/minimizer/gdsq.cc:60-122) to FP32 tolerance;
* registration of synthetically warped frames reduces the cost.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _frames(n, side, seed, dim=2):
    from mia_amd import synth
    if dim == 2:
        return synth.series_frames_2d(n, side, 8.0, 4.0, seed)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.integers(side[0], side[1] + 1, size=3))
        fixed = synth.smoothed_noise(shape, 3.0, int(rng.integers(1 << 30)))
        cs = tuple(synth.spline_csize(s, 8.0, 3) for s in shape[::-1])
        p = synth.random_spline_params(cs, 1.5, 3.0, int(rng.integers(1 << 30)))
        out.append((synth.spline_warp(fixed, cs, p), fixed))
    return out


@pytest.mark.parametrize("dim,n,side", [(2, 7, (40, 120)), (3, 3, (24, 40))])
def test_modes_agree_bitwise(dim, n, side):
    from mia_amd import series
    frames = _frames(n, side, 3, dim)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        res = {m: s.run("lbfgs:maxiter=25", m) for m in (series.BATCHED_GRAPH, series.BATCHED, series.PER_FRAME)}
    p0, c0, e0, _ = res[series.BATCHED_GRAPH]
    for m in (series.BATCHED, series.PER_FRAME):
        p, c, e, _ = res[m]
        assert np.array_equal(p, p0), m
        assert np.array_equal(c, c0), m
        assert np.array_equal(e, e0), m
    assert np.all(e0 <= 25)


def test_registration_reduces_cost():
    import mia_amd
    from mia_amd import series
    frames = _frames(4, (64, 128), 5)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        p, c, e, _ = s.run("lbfgs:maxiter=60", series.BATCHED_GRAPH)
        for i, (mov, fix) in enumerate(frames):
            a, b = series.normalize_pair(mov, fix)
            cs = tuple(mia_amd.spline_csize(v, 8.0, 2) for v in mov.shape[::-1])
            c_init = O.cost(a, b, cs, np.zeros(2 * cs[0] * cs[1]))
            c_fin = O.cost(a, b, cs, p[s.slices[i]])
            assert c_fin == pytest.approx(c[i], rel=1e-5)
            assert c_fin < 0.5 * c_init, (i, c_init, c_fin)


def _gdsq(x, fdf, maxiter, step=0.1, scale=2.0, min_step=1e-6):
    """gdsq.cc:60-122 (xtola = gtola = ftolr = 0)."""
    xwork = x.copy()
    f_old, g = fdf(xwork)
    it = 0
    while it < maxiter and step > min_step:
        it += 1
        xwork = xwork - step * g
        f, gt = fdf(xwork)
        if f < f_old:
            s = -step * g
            aux = f_old - f + s @ gt
            step = step * scale if aux <= 0.0 else 0.5 * np.linalg.norm(s) / aux
            x = xwork.copy()
            g = gt
            f_old = f
        else:
            xwork = x.copy()
            step /= scale
    return x, f_old


def test_gdsq_follows_oracle_trajectory():
    import mia_amd
    from mia_amd import series
    frames = _frames(1, (56, 72), 11)
    mov, fix = frames[0]
    with series.SeriesRegistration(frames, rate=8.0) as s:
        p, c, e, _ = s.run("gdsq:maxiter=12", series.BATCHED_GRAPH)
    a, b = series.normalize_pair(mov, fix)
    cs = tuple(mia_amd.spline_csize(v, 8.0, 2) for v in mov.shape[::-1])
    x_ref, f_ref = _gdsq(np.zeros(2 * cs[0] * cs[1]), lambda q: O.cost_grad(a, b, cs, q), 12)
    assert c[0] == pytest.approx(f_ref, rel=1e-4)
    assert np.abs(p - x_ref).max() <= 1e-3 * max(1.0, np.abs(x_ref).max())
    assert e[0] == 13


def test_bad_minimizer_descriptor():
    from mia_amd import series
    frames = _frames(1, (40, 48), 1)
    with series.SeriesRegistration(frames) as s:
        with pytest.raises(ValueError):
            s.run("lbfgs:maxiter=0")
        with pytest.raises(ValueError):
            s.run("gdsq:scale=1.0")
        with pytest.raises(ValueError):
            s.run("newton:maxiter=3")


@pytest.mark.parametrize("dim,n,side", [(2, 6, (48, 140)), (3, 2, (24, 36))])
def test_device_optimizer_batched_equals_per_frame(dim, n, side):
    """On-device L-BFGS: one graph for all frames and one context per frame
    give bitwise identical parameters, costs and evaluation counts."""
    from mia_amd import series
    frames = _frames(n, side, 21, dim)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        pb, cb, eb, _ = s.run("lbfgs:maxiter=30", series.DEVICE_BATCHED)
        pf, cf, ef, _ = s.run("lbfgs:maxiter=30", series.DEVICE_PER_FRAME)
    assert np.array_equal(pb, pf)
    assert np.array_equal(cb, cf)
    assert np.array_equal(eb, ef)
    assert np.all(eb <= 30)


def test_device_optimizer_both_search_paths():
    """Frames above and below the register-resident limit of the device
    L-BFGS direction search (2048 parameters) in one batch: bitwise equal to
    one context per frame, and close to the host L-BFGS."""
    from mia_amd import series, synth
    frames = synth.series_frames_2d(2, (60, 80), 8.0, 4.0, 31) + synth.series_frames_2d(2, (300, 330), 8.0, 4.0, 32)
    sizes = [2 * mia_amd_csize(f[1].shape) for f in frames]
    assert min(sizes) <= 2048 < max(sizes)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        pb, cb, eb, _ = s.run("lbfgs:maxiter=25", series.DEVICE_BATCHED)
        pf, cf, ef, _ = s.run("lbfgs:maxiter=25", series.DEVICE_PER_FRAME)
        ph, ch, eh, _ = s.run("lbfgs:maxiter=6", series.BATCHED)
        pd, cd, ed, _ = s.run("lbfgs:maxiter=6", series.DEVICE_BATCHED)
    assert np.array_equal(pb, pf)
    assert np.array_equal(cb, cf)
    assert np.array_equal(eb, ef)
    assert np.array_equal(eh, ed)
    np.testing.assert_allclose(cd, ch, rtol=1e-9)


@pytest.mark.parametrize("dim,n,side", [(2, 9, (40, 150)), (3, 3, (24, 40))])
def test_device_split_equals_batched_and_per_frame(dim, n, side):
    """The frames in batches whose iterations overlap on streams of their own
    (mia_hip_optimize_many): bitwise the parameters, costs and evaluation
    counts of one batch and of one context per frame."""
    from mia_amd import series
    frames = _frames(n, side, 23, dim)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        ps, cs, es, _ = s.run("lbfgs:maxiter=30", series.DEVICE_SPLIT)
        pb, cb, eb, _ = s.run("lbfgs:maxiter=30", series.DEVICE_BATCHED)
        pf, cf, ef, _ = s.run("lbfgs:maxiter=30", series.DEVICE_PER_FRAME)
        ps2, cs2, es2, _ = s.run("lbfgs:maxiter=30", series.DEVICE_SPLIT)
    for a, b, c in ((ps, pb, pf), (cs, cb, cf), (es, eb, ef)):
        assert np.array_equal(a, b)
        assert np.array_equal(a, c)
    assert np.array_equal(ps2, ps) and np.array_equal(cs2, cs) and np.array_equal(es2, es)
    assert np.all(es <= 30)


def test_device_split_one_frame():
    """One frame: the split mode runs it as one batch, with the results of the
    batched mode."""
    from mia_amd import series
    frames = _frames(1, (50, 60), 24)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        ps, cs, es, _ = s.run("lbfgs:maxiter=15", series.DEVICE_SPLIT)
        pb, cb, eb, _ = s.run("lbfgs:maxiter=15", series.DEVICE_BATCHED)
    assert np.array_equal(ps, pb) and np.array_equal(cs, cb) and np.array_equal(es, eb)


def test_optimize_many_equals_each_alone():
    """mia_hip_optimize_many over four contexts (the first and the last share
    a stream): a 2D batch with frames on both sides of the one-block step's
    2048-parameter limit, a 3D volume stepped in segments with the divcurl
    penalty attached, and two more 2D batches; each context's parameters,
    costs and evaluation counts are bitwise those of mia_hip_optimize on that
    context alone, for two evaluation budgets."""
    import mia_amd
    from mia_amd import level, series, synth
    fr_a = synth.series_frames_2d(2, (60, 80), 8.0, 4.0, 61) + synth.series_frames_2d(1, (300, 320), 8.0, 4.0, 62)
    fr_b = synth.series_frames_2d(3, (90, 130), 8.0, 4.0, 63)
    fr_c = synth.series_frames_2d(2, (40, 70), 8.0, 4.0, 66)
    shape = (40, 36, 44)
    fix3 = synth.smoothed_noise(shape, 2.0, 64)
    cs3 = tuple(mia_amd.spline_csize(v, 4.0, 3) for v in shape[::-1])
    mov3 = synth.spline_warp(fix3, cs3, synth.random_spline_params(cs3, 1.5, 3.0, 65))
    assert 3 * cs3[0] * cs3[1] * cs3[2] > 2048

    def ctx2d(frames):
        pairs = [series.normalize_pair(m, f) for m, f in frames]
        shapes = [p[0].shape for p in pairs]
        cs = [(mia_amd.spline_csize(v[1], 8.0, 2), mia_amd.spline_csize(v[0], 8.0, 2)) for v in shapes]
        c = mia_amd.HipContext(0)
        c.setup(2, shapes, cs, 3)
        for k, p in enumerate(pairs):
            c.set_images(k, p[0], p[1])
        return c

    ca, cb, cc = ctx2d(fr_a), ctx2d(fr_b), ctx2d(fr_c)
    c3 = mia_amd.HipContext(0)
    pen = level.DivcurlPenalty(cs3, tuple(float(v) for v in shape[::-1]))
    try:
        c3.setup(3, [shape], [cs3], 3)
        c3.set_images(0, mov3, fix3)
        c3.set_penalty(0, pen, 0.01)
        ctxs = [ca, c3, cb, cc]
        x0 = [np.zeros(c.n_params_total) for c in ctxs]
        for maxeval in (25, 8):
            alone = [c.optimize(x, maxeval=maxeval, iters_per_graph=10)[:3] for c, x in zip(ctxs, x0)]
            joint, sec = mia_amd.HipContext.optimize_many(ctxs, x0, maxeval=maxeval, iters_per_graph=10)
            assert sec > 0.0
            for k in range(4):
                for a, b in zip(joint[k], alone[k]):
                    assert np.array_equal(a, b), (maxeval, k)
                assert np.all(joint[k][2] <= maxeval)
        # the contexts keep working on their own streams afterwards
        again = ca.optimize(x0[0], maxeval=25, iters_per_graph=10)[:3]
        first = mia_amd.HipContext.optimize_many([ca], [x0[0]], maxeval=25)[0][0]
        for a, b in zip(again, first):
            assert np.array_equal(a, b)
        with pytest.raises(ValueError):
            mia_amd.HipContext.optimize_many([ca, ca], [x0[0], x0[0]], maxeval=5)
        # a context that is not ready fails the call before any iteration, and
        # the others keep their own streams and results
        with mia_amd.HipContext(0) as bare:
            bare.setup(2, [(40, 48)], [(mia_amd.spline_csize(48, 8.0, 2), mia_amd.spline_csize(40, 8.0, 2))], 3)
            with pytest.raises(mia_amd._lib.MiaHipError):
                mia_amd.HipContext.optimize_many([cb, bare], [x0[2], np.zeros(bare.n_params_total)], maxeval=5)
        again_b = cb.optimize(x0[2], maxeval=8, iters_per_graph=10)[:3]
        for a, b in zip(again_b, alone[2]):
            assert np.array_equal(a, b)
    finally:
        c3.set_penalty(0, None, 0.0)
        for c in (ca, cb, cc, c3):
            c.close()
        pen.close()


def test_device_lbfgs_ring_wraparound():
    """memory = 2: from the third accepted step on, every step drops the
    oldest pair and overwrites its ring slot (and its rows of the pair-product
    tables), so twelve evaluations wrap the ring several times.  Host and
    device L-BFGS (same algorithm, different summation order) must agree
    tightly on frames on both sides of the 2048-parameter limit of the
    register-resident direction."""
    from mia_amd import series, synth
    frames = synth.series_frames_2d(2, (60, 80), 8.0, 4.0, 41) + synth.series_frames_2d(2, (300, 330), 8.0, 4.0, 42)
    sizes = [2 * mia_amd_csize(f[1].shape) for f in frames]
    assert min(sizes) <= 2048 < max(sizes)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        ph, ch, eh, _ = s.run("lbfgs:maxiter=12,memory=2", series.BATCHED)
        pd, cd, ed, _ = s.run("lbfgs:maxiter=12,memory=2", series.DEVICE_BATCHED)
    print("evaluations", eh, ed, "cost rel", np.abs(cd - ch) / np.abs(ch),
          "param max abs", np.abs(pd - ph).max())
    assert np.array_equal(eh, ed)
    assert np.all(eh >= 10)
    np.testing.assert_allclose(cd, ch, rtol=1e-9)
    assert np.abs(pd - ph).max() <= 1e-7 * max(1.0, np.abs(ph).max())


def mia_amd_csize(shape):
    import mia_amd
    return mia_amd.spline_csize(shape[1], 8.0, 2) * mia_amd.spline_csize(shape[0], 8.0, 2)


def test_device_optimizer_follows_host_lbfgs():
    """Same algorithm on the device and on the host (different summation
    order of the dot products): the first iterations agree tightly and the
    final costs of a longer run agree closely."""
    from mia_amd import series
    frames = _frames(3, (60, 110), 8)
    with series.SeriesRegistration(frames, rate=8.0) as s:
        ph, ch, eh, _ = s.run("lbfgs:maxiter=6", series.BATCHED)
        pd, cd, ed, _ = s.run("lbfgs:maxiter=6", series.DEVICE_BATCHED)
        assert np.array_equal(eh, ed)
        np.testing.assert_allclose(cd, ch, rtol=1e-9)
        assert np.abs(pd - ph).max() <= 1e-7 * max(1.0, np.abs(ph).max())
        _, ch2, _, _ = s.run("lbfgs:maxiter=60", series.BATCHED)
        _, cd2, _, _ = s.run("lbfgs:maxiter=60", series.DEVICE_BATCHED)
    np.testing.assert_allclose(cd2, ch2, rtol=2e-2)
    assert np.all(cd2 < cd)


def test_device_optimizer_rejects_gdsq():
    from mia_amd import series
    frames = _frames(1, (40, 48), 2)
    with series.SeriesRegistration(frames) as s:
        with pytest.raises(ValueError):
            s.run("gdsq:maxiter=5", series.DEVICE_BATCHED)


@pytest.mark.parametrize("ipg", [1, 7, 10])
def test_optimize_stream_and_graph_launches_agree(ipg):
    """mia_hip_optimize queued as direct stream launches (next chunk queued
    ahead) and as one captured graph per chunk: the same kernels in the same
    order, so parameters, costs and evaluation counts are bitwise equal, for
    chunk sizes that divide maxeval and that do not, with frames on both sides
    of the 2048-parameter limit of the one-block step."""
    import mia_amd
    from mia_amd import series, synth
    frames = synth.series_frames_2d(2, (60, 80), 8.0, 4.0, 51) + synth.series_frames_2d(1, (300, 320), 8.0, 4.0, 52)
    pairs = [series.normalize_pair(m, f) for m, f in frames]
    shapes = [p[0].shape for p in pairs]
    cs = [(mia_amd.spline_csize(s[1], 8.0, 2), mia_amd.spline_csize(s[0], 8.0, 2)) for s in shapes]
    out = {}
    with mia_amd.HipContext(0) as c:
        c.setup(2, shapes, cs, 3)
        for k, p in enumerate(pairs):
            c.set_images(k, p[0], p[1])
        with pytest.raises(ValueError):
            mia_amd._lib.check(c._L.mia_hip_set_optimize_launch(c._h, 2))
        with pytest.raises(ValueError):
            c.set_optimize_launch("persistent")
        for mode in ("graph", "stream"):
            c.set_optimize_launch(mode)
            out[mode] = c.optimize(np.zeros(c.n_params_total), maxeval=20, iters_per_graph=ipg)[:3]
    for a, b in zip(out["stream"], out["graph"]):
        assert np.array_equal(a, b)
    assert np.all(out["stream"][2] <= 20)
    ref = out["stream"]
    with mia_amd.HipContext(0) as c:
        c.setup(2, shapes, cs, 3)
        for k, p in enumerate(pairs):
            c.set_images(k, p[0], p[1])
        again = c.optimize(np.zeros(c.n_params_total), maxeval=20, iters_per_graph=20)[:3]
    for a, b in zip(ref, again):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("memory", [10, 14])
def test_device_lbfgs_memory_paths_follow_host(memory):
    """The one-block device step holds up to 10 stored pairs in registers and
    reads more from memory: both paths against the host L-BFGS (same
    algorithm, other summation order) over enough evaluations to store more
    than 10 pairs, on 2D frames (vectors moved in adjacent pairs) and on 3D
    volumes of an odd parameter count (one element per thread and slot)."""
    from mia_amd import series
    frames2 = _frames(3, (60, 110), 61)
    frames3 = _frames(2, (22, 26), 65, 3)
    sizes3 = [3 * np.prod([mia_amd_csize3(s) for s in f[1].shape]) for f in frames3]
    assert any(n % 2 == 1 for n in sizes3)
    for frames in (frames2, frames3):
        with series.SeriesRegistration(frames, rate=8.0) as s:
            desc = "lbfgs:maxiter=28,memory=%d" % memory
            ph, ch, eh, _ = s.run(desc, series.BATCHED)
            pd, cd, ed, _ = s.run(desc, series.DEVICE_BATCHED)
        print("memory", memory, "evaluations", eh, ed, "cost rel", np.abs(cd - ch) / np.abs(ch))
        assert np.array_equal(eh, ed)
        np.testing.assert_allclose(cd, ch, rtol=1e-7)
        assert np.abs(pd - ph).max() <= 1e-6 * max(1.0, np.abs(ph).max())


def mia_amd_csize3(n):
    import mia_amd
    return mia_amd.spline_csize(n, 8.0, 3)
