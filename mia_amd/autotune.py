"""Autotuner of the 3D warp kernel's tile shapes (include/mia_hip.h).

    python -m mia_amd.autotune [--out PATH] [--degree 0|1|3] [--buckets 15 24]

times every tile variant on synthetic instances of 2^b voxels per size
bucket b and interpolation degree, writes the fastest per bucket as JSON
keyed by the GPU architecture (default mia_amd/tuning/warp3d_<arch>.json,
where the library looks for it) and prints the table.
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

from . import _lib
from ._lib import check

_HERE = os.path.dirname(os.path.abspath(__file__))


def variants():
    L = _lib.load()
    dims = (ctypes.c_int * 48)()
    n = ctypes.c_int()
    check(L.mia_hip_warp_variants(dims, 16, ctypes.byref(n)))
    return [tuple(dims[3 * v:3 * v + 3]) for v in range(n.value)]


def device_arch(device=0):
    import torch
    return torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]


def run(out_path=None, degree=0, b_min=15, b_max=24, device=0):
    """Tune and return (table dict, microseconds array [deg][bucket][variant])."""
    L = _lib.load()
    nv = len(variants())
    nd = 2 if degree == 0 else 1
    nb = b_max - b_min + 1
    us = np.zeros(nd * nb * nv)
    path = out_path.encode() if out_path else None
    check(L.mia_hip_autotune_warp(int(device), int(degree), int(b_min), int(b_max), path, us.ctypes.data))
    table = json.load(open(out_path)) if out_path else None
    return table, us.reshape(nd, nb, nv)


def tuning_source():
    buf = ctypes.create_string_buffer(1024)
    n = _lib.load().mia_hip_warp_tuning_source(buf, 1024)
    return buf.value.decode(), n


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=None)
    ap.add_argument("--degree", type=int, default=0, choices=(0, 1, 3))
    ap.add_argument("--buckets", type=int, nargs=2, default=(15, 24))
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    out = a.out or os.path.join(_HERE, "tuning", "warp3d_%s.json" % device_arch(a.device))
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    table, us = run(out, a.degree, a.buckets[0], a.buckets[1], a.device)
    v = variants()
    print("variants (tile x, y, z):", v)
    for row in table["table"]:
        print("degree %d  bucket 2^%d (%s^3)  best %s  us %s" % (row["degree"], row["bucket"], row["size"][0],
                                                               v[row["variant"]], row["us"]))
    print("written", out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
