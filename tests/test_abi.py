"""CPU tests of the C-ABI library: it loads, exports every declared symbol,
reports the absence of a GPU correctly, and its host-side generator matches
the Python restatement.  No compute call needs a GPU here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("zkmetal.h", "zkmsm.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:zkmetal|zkmsm)_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(zk.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_msm_subset_is_declared():
    # the MSM subset of Sources/zkMetal-ffi/include/zkmetal.h
    want = ["zkmetal_msm_engine_create", "zkmetal_msm_engine_destroy", "zkmetal_bn254_msm",
            "zkmetal_bn254_msm_auto", "zkmetal_bn254_msm_u8", "zkmetal_bn254_msm_u16",
            "zkmetal_bn254_msm_u32", "zkmetal_bn254_msm_u8_auto", "zkmetal_bn254_msm_u16_auto",
            "zkmetal_bn254_msm_u32_auto", "zkmetal_bn254_msm_batch", "zkmetal_bn254_msm_batch_auto",
            "zkmetal_set_shader_dir", "zkmetal_gpu_available", "zkmetal_version"]
    assert set(want) <= set(declared_functions())


def test_headers_compile_as_c_and_cxx(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "zkmsm.h"\nint main(void){ZkMetalMSMEngine e=0; (void)e; '
                   'return zkmetal_gpu_available() < 0;}\n')
    inc = os.path.join(ROOT, "include")
    for cc, ext in (("gcc", "c"), ("g++", "cpp")):
        s = tmp_path / f"use.{ext}"
        s.write_text(src.read_text())
        out = tmp_path / f"a_{ext}"
        r = subprocess.run([cc, "-Wall", "-Werror", "-I", inc, str(s), "-o", str(out),
                            zk.LIB_PATH, f"-Wl,-rpath,{os.path.dirname(zk.LIB_PATH)}"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_version_and_shader_dir():
    assert zk.version().startswith("0.2.0")
    zk.lib().zkmetal_set_shader_dir(b"/nonexistent")
    zk.lib().zkmetal_set_shader_dir(None)


def test_null_engine_is_invalid_input():
    L = zk.lib()
    buf = (ctypes.c_uint8 * 96)()
    assert L.zkmetal_bn254_msm(None, buf, buf, 1, buf, buf, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmetal_bn254_msm_batch(None, buf, buf, buf, 1, buf) == zk.ERR_INVALID_INPUT
    # the sibling-curve entry points reject a null engine the same way
    assert L.zkmsm_grumpkin_msm(None, buf, buf, 1, buf, buf, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_grumpkin_msm_batch(None, buf, buf, buf, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_grumpkin_msm_batch_device(None, buf, buf, buf, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_grumpkin_points_generate(None, 1, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_bn254_g2_msm(None, buf, buf, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_bn254_g2_msm_batch(None, buf, buf, buf, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_bn254_g2_msm_batch_device(None, buf, buf, buf, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_bn254_g2_points_generate(None, 1, 1, buf) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_test_g2(None, 0, buf, buf, buf, 1) == zk.ERR_INVALID_INPUT
    # and so do the tuning tables
    tab = (ctypes.c_uint32 * 33)()
    assert L.zkmsm_set_window_table(None, tab, 33) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_set_batch_window_table(None, tab, 33) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_set_schedule_table(None, 0, tab, tab, 33) == zk.ERR_INVALID_INPUT
    assert L.zkmsm_set_schedule_table(None, 1, None, None, 0) == zk.ERR_INVALID_INPUT
    L.zkmetal_msm_engine_destroy(None)


@pytest.mark.skipif(zk.gpu_available(), reason="checks the no-GPU path")
def test_no_gpu_is_reported():
    h = ctypes.c_void_p()
    assert zk.lib().zkmetal_msm_engine_create(ctypes.byref(h)) == zk.ERR_NO_GPU
    buf = (ctypes.c_uint8 * 96)()
    assert zk.lib().zkmetal_bn254_msm_auto(buf, buf, 1, buf, buf, buf) == zk.ERR_NO_GPU
    with pytest.raises(zk.MSMError):
        zk.MetalMSM()


@pytest.mark.parametrize("dist", [0, 1, 2])
def test_generator_matches_python(dist):
    n = 600
    got = zk.gen_scalars(0xABCDEF, n, dist)
    for i in range(n):
        assert int.from_bytes(got[i].tobytes(), "little") == B.gen_scalar(0xABCDEF, i, dist)


def test_generator_distributions():
    n = 20000
    s = [int.from_bytes(r.tobytes(), "little") for r in zk.gen_scalars(5, n, 1)]
    zeros = sum(1 for x in s if x == 0) / n
    small = sum(1 for x in s if 1 < x < 2**16) / n
    ones = sum(1 for x in s if x in (1, B.R_ORDER - 1)) / n
    assert 0.37 < zeros < 0.43 and 0.27 < small < 0.33 and 0.08 < ones < 0.12
    u = [int.from_bytes(r.tobytes(), "little") for r in zk.gen_scalars(6, 2000, 0)]
    assert all(x < B.R_ORDER for x in u) and max(u).bit_length() == 254


def test_default_window_bits_policy():
    ws = [zk.default_window_bits(1 << k) for k in range(0, 24)]
    assert all(2 <= c <= 16 for c in ws)
    assert ws == sorted(ws)  # non-decreasing with n
    assert zk.default_window_bits(1 << 22) == 16


def test_server_footprint_monotone():
    f = [zk.server_footprint(1 << k) for k in range(0, 24)]
    assert all(a <= b for a, b in zip(f, f[1:]))
    assert f[22] > 64 * (1 << 22)  # at least the scalars and digit arrays
