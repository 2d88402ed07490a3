"""Whole registrations through the multiresolution driver (libmia_host).

Plain 2D registration (src/2dimageregistration.cc:46-106, one level): a
128x128 Gaussian-blob image against its own B-spline warp (knot spacing 16,
coefficients N(0, 1.5) clipped to 3 px), `spline:rate=16`, ssd, L-BFGS with
50 evaluations.  The returned cost must be the FP64 oracle's cost at the
returned parameters on the normalised pair (mia/2d/fullcost/image.cc:81-105),
and the registration must remove most of the misalignment.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_plain_2d_registration():
    from mia_amd import series, synth
    mov, fix, cs, _ = synth.blob_pair_2d()
    assert cs == (10, 10)
    params, grid, stats, final = series.nonrigid_register(mov, fix, "spline:rate=16", "image:cost=[ssd]",
                                                          "lbfgs:maxiter=50", levels=1)
    assert tuple(grid[:2]) == cs and params.size == 200
    assert len(stats) == 1 and 1 < stats[0]["evaluations"] <= 50
    a, b = series.normalize_pair(mov, fix)
    c_id = O.cost(a, b, cs, np.zeros(200))
    c_fin = O.cost(a, b, cs, params)
    print("identity %.6g, final %.6g (oracle %.6g), %d evaluations"
          % (c_id, final, c_fin, stats[0]["evaluations"]))
    assert final == pytest.approx(c_fin, rel=1e-5)
    assert c_fin < 0.2 * c_id
    # the on-device L-BFGS on the same problem reaches a comparable cost
    with series.SeriesRegistration([(mov, fix)], rate=16.0) as s:
        p_d, c_d, e_d, _ = s.run("lbfgs:maxiter=50", series.DEVICE_BATCHED)
    assert c_d[0] == pytest.approx(O.cost(a, b, cs, p_d), rel=1e-5)
    assert c_d[0] < 0.2 * c_id
