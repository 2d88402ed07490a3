"""CPU checks of the host layer: the minimizers against problems with known
answers (compiled C++ test), and the host library's exported symbols."""
import ctypes

import pytest
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_minimizers_cpp(tmp_path):
    exe = tmp_path / "test_minimizer"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", str(exe),
                           os.path.join(ROOT, "tests", "cpp", "test_minimizer.cc")])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")


def test_host_library_exports():
    from mia_amd import series
    text = re.sub(r"/\*.*?\*/", "", open(series.HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(mia_host_\w+)\s*\(", text)))
    assert names == sorted(series.SIGNATURES)
    L = ctypes.CDLL(series.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_normalize_pair_moments():
    import numpy as np
    from mia_amd.series import normalize_pair
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 100, (20, 30)).astype(np.float32)
    b = rng.uniform(50, 300, (20, 30)).astype(np.float32)
    na, nb = normalize_pair(a, b)
    both = np.concatenate([na.ravel(), nb.ravel()]).astype(np.float64)
    assert abs(both.mean()) < 1e-5
    assert abs(both.std(ddof=1) - 1.0) < 1e-5


def _cpp_bin(name):
    b = os.path.join(ROOT, "tests", "cpp", "bin", name)
    if not os.path.exists(b):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), b])
    return b


def test_nonrigid_host_cpp():
    """C++ mirror of the reference's classes, host-side cases
    (tests/cpp/test_nonrigid_host.cc), with the plugin module that overrides
    the library's `ssd` plugin by priority."""
    exe = _cpp_bin("test_nonrigid_host")
    mod = _cpp_bin("test_plugin_module.so")
    env = dict(os.environ, MIA_TEST_PLUGIN_MODULE=mod)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")


def test_transform_file_round_trip(tmp_path):
    """Transformation files in MIA's xml layout: parameters survive exactly
    (17 significant digits), the grid is re-created from creator and size,
    and malformed or inconsistent files raise."""
    import numpy as np
    from mia_amd import series
    rng = np.random.default_rng(4)
    for shape, creator in (((40, 50, 60), "spline:rate=8,penalty=[divcurl:weight=0.01]"),
                           ((96, 80), "spline:rate=16,imgkernel=[bspline:d=1]")):
        dim = len(shape)
        import mia_amd
        cs = [mia_amd.spline_csize(n, 8.0 if dim == 3 else 16.0, dim) for n in shape[::-1]]
        p = rng.normal(size=dim * int(np.prod(cs))).astype(np.float32).astype(np.float64)
        f = tmp_path / ("t%d.x%ddt" % (dim, dim))
        series.save_transform(f, creator, shape, p)
        text = f.read_text()
        assert "<creator>" + creator + "</creator>" in text and "<params>" in text
        cr, sh, grid, q = series.load_transform(f, dim)
        assert cr == creator and sh == shape and list(grid[:dim]) == cs
        assert np.array_equal(q, p)
    with pytest.raises(ValueError):
        series.save_transform(tmp_path / "bad.x3dt", "spline:rate=8", (40, 50, 60), np.zeros(5))
    (tmp_path / "junk.x3dt").write_text("<boost_serialization></boost_serialization>")
    with pytest.raises(RuntimeError):
        series.load_transform(tmp_path / "junk.x3dt", 3)
    with pytest.raises(RuntimeError):
        series.load_transform(tmp_path / "t2.x2dt", 3)   # a 2D file has no size_z
    t = re.sub(r"<params>\n\t<count>\d+", "<params>\n\t<count>7", (tmp_path / "t3.x3dt").read_text())
    (tmp_path / "cut.x3dt").write_text(t)
    with pytest.raises(RuntimeError):
        series.load_transform(tmp_path / "cut.x3dt", 3)


def test_transform_file_attributes(tmp_path):
    """Attributes are written as the string map of
    mia/3d/transio/serialization.hh:43-49 (count, item_version, first/second
    items, XML-escaped) and restored on load (:66-77); malformed entities
    raise instead of looping."""
    import numpy as np
    import mia_amd
    from mia_amd import series
    shape = (30, 34, 38)
    cs = [mia_amd.spline_csize(n, 8.0, 3) for n in shape[::-1]]
    p = np.linspace(-1, 1, 3 * int(np.prod(cs))).astype(np.float32).astype(np.float64)
    attr = {"affine": "1 0 0 0 1 0 0 0 1", "note": "a<b & \"c\" 'd'>", "z": ""}
    f = tmp_path / "a.x3dt"
    series.save_transform(f, "spline:rate=8", shape, p, attr)
    text = f.read_text()
    assert "<count>3</count>" in text
    assert "<first>note</first>" in text and "a&lt;b &amp; &quot;c&quot; &apos;d&apos;&gt;" in text
    cr, sh, grid, q, back = series.load_transform(f, 3, with_attributes=True)
    assert back == attr and np.array_equal(q, p)
    # no attributes: an empty map, as before
    series.save_transform(tmp_path / "e.x3dt", "spline:rate=8", shape, p)
    assert series.load_transform(tmp_path / "e.x3dt", 3, with_attributes=True)[4] == {}
    # '&' without ';' (once looped forever), an unknown entity, a bad count
    for bad in (text.replace("&amp;", "&x", 1), text.replace("&amp;", "&bogus;", 1),
                text.replace("<count>3</count>", "<count>4</count>", 1)):
        (tmp_path / "bad.x3dt").write_text(bad)
        with pytest.raises(RuntimeError):
            series.load_transform(tmp_path / "bad.x3dt", 3)
    (tmp_path / "c.x3dt").write_text(text.replace("<creator>spline:rate=8", "<creator>spline:rate=8&x"))
    with pytest.raises(RuntimeError):
        series.load_transform(tmp_path / "c.x3dt", 3)
