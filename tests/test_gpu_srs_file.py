"""GPU tests of the SRS-file loader (zkmsm_srs_load_ptau[_file]; the
reference's .ptau layout, Sources/zkMetal/KZG/SRSFileFormat.swift:57-160):
loaded bases equal the points written, MSMs over them equal MSMs over the same
bases uploaded directly, and malformed files or invalid points are refused."""
import struct

import numpy as np
import pytest

import zkmetal_amd as zk
from helpers import affine_bytes_to_points
from oracle import bn254 as B

pytestmark = pytest.mark.gpu

SEED = 0x5A5A
N = 3000


def ptau_bytes(points, extra_sections=True):
    """points: list of affine (x, y) integers or None -> the file image"""
    g1 = bytearray()
    for p in points:
        x, y = (0, 0) if p is None else p
        g1 += x.to_bytes(32, "big") + y.to_bytes(32, "big")
    secs = [(1, struct.pack("<III", 1, 12, 12))]
    if extra_sections:
        secs.append((7, b"contributions"))
    secs.append((2, bytes(g1)))
    if extra_sections:
        secs.append((3, bytes(256)))  # two G2 points (unused by the loader)
    out = bytearray(b"ptau" + struct.pack("<II", 1, len(secs)))
    for t, d in secs:
        out += struct.pack("<IQ", t, len(d)) + d
    return bytes(out)


@pytest.fixture(scope="module")
def ref_points():
    e = zk.MetalMSM()
    e.srs_generate(SEED, N)
    raw = e.srs_download(N)
    e.close()
    return raw, affine_bytes_to_points(bytes(raw))


def test_load_matches_points(ref_points, tmp_path):
    raw, pts = ref_points
    pts = list(pts)
    pts[5] = None  # an identity base
    img = ptau_bytes(pts)
    e = zk.MetalMSM()
    try:
        assert e.srs_load_ptau(img) == N
        back = affine_bytes_to_points(bytes(e.srs_download(N)))
        assert back == pts
        # MSMs over the loaded table == MSMs over the same bases given directly
        sc = zk.gen_scalars(SEED + 1, N, 1)
        direct = bytearray(raw)
        direct[5 * 64:6 * 64] = bytes(64)
        want = e.msm(bytes(direct), sc.tobytes())
        assert e.msm_batch_resident(sc, [N])[0] == want
        # from a path, a prefix only
        f = tmp_path / "srs.ptau"
        f.write_bytes(ptau_bytes(pts, extra_sections=False))
        assert e.srs_load_ptau(path=str(f), max_points=1000) == 1000
        assert e.msm_batch_resident(sc[:1000], [1000])[0] == e.msm(bytes(direct[:64000]), sc[:1000].tobytes())
    finally:
        e.close()


def test_invalid_files_and_points(ref_points):
    _, pts = ref_points
    pts = list(pts[:64])
    e = zk.MetalMSM()
    try:
        good = ptau_bytes(pts)
        assert e.srs_load_ptau(good) == 64
        bad_inputs = [
            b"ptaX" + good[4:],                      # magic
            good[:4] + struct.pack("<I", 2) + good[8:],  # version
            good[:40],                               # truncated section
            ptau_bytes([]),                          # no G1 points
        ]
        off = list(pts)
        off[3] = (off[3][0], (off[3][1] + 1) % B.P)  # not on the curve
        bad_inputs.append(ptau_bytes(off))
        big = list(pts)
        big[7] = (big[7][0] + B.P, big[7][1]) if big[7][0] + B.P < 2**256 else big[7]
        bad_inputs.append(ptau_bytes(big))           # x >= p (same residue)
        hdr = bytearray(good)
        struct.pack_into("<I", hdr, 24, 2)          # curve id 2 (BLS12-381)
        bad_inputs.append(bytes(hdr))
        for img in bad_inputs:
            with pytest.raises(zk.MSMError):
                e.srs_load_ptau(img)
        with pytest.raises(zk.MSMError):
            e.srs_load_ptau(path="/nonexistent/srs.ptau")
    finally:
        e.close()
