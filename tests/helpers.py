"""Shared helpers for the tests (canonical affine comparison, fixtures)."""
import glob
import json
import os

import numpy as np

from oracle import bn254 as B

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_msm_fixtures():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "msm_*.json"))):
        with open(path) as fh:
            out.append(json.load(fh))
    return out


def expected_affine(fx):
    e = fx["expected"]
    return None if e["inf"] else (int(e["x"], 16), int(e["y"], 16))


def result_affine(r96):
    """96-byte Jacobian Montgomery result -> canonical affine (None = infinity)."""
    r96 = bytes(r96)
    return B.jacobian_bytes_to_affine(r96[:32], r96[32:64], r96[64:96])


def assert_normalised(r96):
    """The library returns Z = 1 (Montgomery) or the identity (1, 1, 0)."""
    r96 = bytes(r96)
    one = B.to_mont(1).to_bytes(32, "little")
    if r96[64:] == bytes(32):
        assert r96[:32] == one and r96[32:64] == one
    else:
        assert r96[64:] == one


def scalars_to_ints(arr):
    a = np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1, 32)
    return [int.from_bytes(a[i].tobytes(), "little") for i in range(a.shape[0])]


def affine_bytes_to_points(buf):
    b = bytes(buf)
    pts = []
    for i in range(len(b) // 64):
        x = int.from_bytes(b[64 * i:64 * i + 32], "little")
        y = int.from_bytes(b[64 * i + 32:64 * i + 64], "little")
        pts.append(None if x == 0 and y == 0 else (B.from_mont(x), B.from_mont(y)))
    return pts


def closed_form(ks, ss):
    """(sum s_i k_i mod r) * G from integer lists."""
    acc = 0
    for s, k in zip(ss, ks):
        acc += s * k
    return B.mul(B.G, acc % B.R_ORDER)


def config_fixture():
    """tests/golden/configs.json: seeds and expected affine results of the
    benchmarked configurations."""
    with open(os.path.join(GOLDEN, "configs.json")) as fh:
        return json.load(fh)


def config_expected_point(case):
    """Expected affine point of one configs.json case (None = identity); G2
    coordinates as (c0, c1) pairs."""
    e = case["expected"]
    if e["inf"]:
        return None
    if case["curve"] == "g2":
        return (tuple(int(v, 16) for v in e["x"]), tuple(int(v, 16) for v in e["y"]))
    return (int(e["x"], 16), int(e["y"], 16))
