"""Autotuner of the 3D and 2D warp tiles and the 3D stage-1 layout
(include/mia_hip.h, mia_hip_autotune).

    python -m mia_amd.autotune [--out PATH] [--kinds warp3d warp2d stage1_3d]
                               [--degree 0|1|3] [--buckets3d 15 24]
                               [--buckets2d 12 18] [--buckets-stage1 15 24]

times every choice of each kind on synthetic instances of 2^b voxels (3D
cubes) or pixels (batches of 16 2D squares) per size bucket b and
interpolation degree, writes the fastest per bucket into one JSON table keyed
by the GPU architecture (default mia_amd/tuning/<arch>.json, where the library
looks for it) and prints it.  The architecture comes from the library
(mia_hip_device_arch).
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

KINDS = ("warp3d", "warp2d", "stage1_3d")
STAGE1_LAYOUTS = ("slice", "zcell")


def variants():
    """3D warp tile shapes (x, y, z) by variant id."""
    L = _lib.load()
    dims = (ctypes.c_int * 48)()
    n = ctypes.c_int()
    check(L.mia_hip_warp_variants(dims, 16, ctypes.byref(n)))
    return [tuple(dims[3 * v:3 * v + 3]) for v in range(n.value)]


def variants_2d():
    """2D warp tile shapes (x, y, "box" or "direct") by variant id."""
    L = _lib.load()
    dims = (ctypes.c_int * 48)()
    n = ctypes.c_int()
    check(L.mia_hip_warp2d_variants(dims, 16, ctypes.byref(n)))
    return [(dims[3 * v], dims[3 * v + 1], "box" if dims[3 * v + 2] else "direct") for v in range(n.value)]


def choices(kind):
    return {"warp3d": variants, "warp2d": variants_2d, "stage1_3d": lambda: list(STAGE1_LAYOUTS)}[kind]()


def device_arch(device=0):
    """The GPU architecture name the tuning table is keyed by ("gfx950")."""
    buf = ctypes.create_string_buffer(64)
    check(_lib.load().mia_hip_device_arch(int(device), buf, 64))
    return buf.value.decode()


def default_path(device=0):
    return os.path.join(_HERE, "tuning", "%s.json" % device_arch(device))


def run(out_path=None, degree=0, b_min=15, b_max=24, device=0, kind="warp3d"):
    """Tune one kind and return (table dict or None, microseconds array
    [degree][bucket][choice]).  Entries already in out_path (same
    architecture) for other kinds, degrees or buckets are kept."""
    L = _lib.load()
    k = KINDS.index(kind)
    nc = len(choices(kind))
    nd = 2 if degree == 0 else 1
    nb = b_max - b_min + 1
    us = np.zeros(nd * nb * nc)
    path = out_path.encode() if out_path else None
    check(L.mia_hip_autotune(int(device), k, int(degree), int(b_min), int(b_max), path, us.ctypes.data))
    table = json.load(open(out_path)) if out_path else None
    return table, us.reshape(nd, nb, nc)


def tuning_source():
    buf = ctypes.create_string_buffer(1024)
    n = _lib.load().mia_hip_warp_tuning_source(buf, 1024)
    return buf.value.decode(), n


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=None)
    ap.add_argument("--kinds", nargs="+", default=list(KINDS), choices=KINDS)
    ap.add_argument("--degree", type=int, default=0, choices=(0, 1, 3))
    ap.add_argument("--buckets3d", type=int, nargs=2, default=(15, 24))
    ap.add_argument("--buckets2d", type=int, nargs=2, default=(12, 18))
    ap.add_argument("--buckets-stage1", type=int, nargs=2, default=(15, 24))
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    out = a.out or default_path(a.device)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    buckets = {"warp3d": a.buckets3d, "warp2d": a.buckets2d, "stage1_3d": a.buckets_stage1}
    table = None
    for kind in a.kinds:
        table, _ = run(out, a.degree, buckets[kind][0], buckets[kind][1], a.device, kind)
        print("tuned", kind, flush=True)
    for row in table["table"]:
        kind = row.get("kernel", "warp3d")
        print("%-9s degree %d  bucket 2^%d %s  best %s  us %s" % (kind, row["degree"], row["bucket"], row.get("size"),
                                                                  choices(kind)[row["variant"]], row["us"]))
    print("written", out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
