"""The batched 3D workload (16 volumes, sides U[64,160], rate 6) against the
FP64 oracle: cost, gradient and warped image of every volume, evaluated in
one batched context.  Prints the errors as one JSON line."""
import json
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batched_volumes_vs_oracle():
    sys.path.insert(0, ROOT)
    import bench
    import mia_amd
    data = bench.batch_volumes_3d(16)
    shapes = [d[0].shape for d in data]
    with mia_amd.HipContext(0) as c:
        c.setup(3, shapes, [d[2] for d in data], 3)
        for i, d in enumerate(data):
            c.set_images(i, d[0], d[1])
        costs, grad = c.cost_grad(np.concatenate([d[3] for d in data]))
        warped = [c.warp(i, d[3]) for i, d in enumerate(data)]
        variants = [c.warp_variant(i) for i in range(len(data))]
    rows = []
    off = 0
    for i, (mov, fix, cs, p) in enumerate(data):
        c_ref, g_ref, w_ref = O.cost_grad(mov, fix, cs, p, want_warped=True)
        g = grad[off:off + p.size]
        off += p.size
        rows.append({"shape": list(shapes[i]), "variant": variants[i],
                     "cost_rel": float(abs(costs[i] - c_ref) / abs(c_ref)),
                     "grad_rel_inf": float(np.abs(g - g_ref).max() / np.abs(g_ref).max()),
                     "warped_max_abs": float(np.abs(warped[i] - w_ref).max())})
    print(json.dumps({"batched_3d_oracle": rows,
                      "max_cost_rel": max(r["cost_rel"] for r in rows),
                      "max_grad_rel_inf": max(r["grad_rel_inf"] for r in rows),
                      "max_warped_abs": max(r["warped_max_abs"] for r in rows)}))
    for r in rows:
        assert r["cost_rel"] < 1e-5, r
        assert r["grad_rel_inf"] < 1e-5, r
        assert r["warped_max_abs"] < 2e-5, r


def test_mixed_stage1_paths_bitwise():
    """A batch whose instances take different stage-1 paths (x rows longer
    than a block, y bands taller than the register window, and instances that
    fit both) matches one context per instance bit for bit, and the oracle."""
    import mia_amd
    from mia_amd import synth
    shapes = [(12, 20, 300), (25, 38, 27), (37, 25, 27), (30, 41, 64)]
    data = []
    for k, sh in enumerate(shapes):
        fix = synth.smoothed_noise(sh, 2.0, 500 + k)
        mov = synth.smoothed_noise(sh, 2.0, 600 + k)
        cs = tuple(O.spline_csize(n, 8.0, 3) for n in sh[::-1])
        data.append((mov, fix, cs, synth.random_spline_params(cs, 1.5, 3.0, 700 + k)))
    with mia_amd.HipContext(0) as c:
        c.setup(3, shapes, [d[2] for d in data], 3)
        print("stage-1 paths", [c.stage1_path(i) for i in range(len(shapes))], flush=True)
        assert c.stage1_path(0) == 0 and c.stage1_path(1) == 0
        for i, d in enumerate(data):
            c.set_images(i, d[0], d[1])
        cb, gb = c.cost_grad(np.concatenate([d[3] for d in data]))
    off = 0
    for i, (mov, fix, cs, p) in enumerate(data):
        with mia_amd.HipContext(0) as c1:
            c1.setup(3, [shapes[i]], [cs], 3)
            c1.set_images(0, mov, fix)
            c1c, g1 = c1.cost_grad(p)
        assert c1c[0] == cb[i], i
        assert np.array_equal(g1, gb[off:off + p.size]), i
        off += p.size
        c_ref, g_ref = O.cost_grad(mov, fix, cs, p)
        assert c1c[0] == pytest.approx(c_ref, rel=1e-5)
        assert np.abs(g1 - g_ref).max() <= 1e-5 * np.abs(g_ref).max()
