"""CPU checks of the host layer: the minimizers against problems with known
answers (compiled C++ test), and the host library's exported symbols."""
import ctypes
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
