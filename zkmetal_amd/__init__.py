"""zkmetal_amd -- BN254 G1 multi-scalar multiplication on AMD MI355X (gfx950).

Python view of libzkmetal_amd.so (the C ABI in include/zkmetal.h and
include/zkmsm.h).  `MetalMSM` mirrors the reference engine API
(Sources/zkMetal/MSM/MSMEngine.swift:17,287: `MetalMSM().msm(points:scalars:)`)
with the same argument meaning and error behaviour (MSMError.invalidInput for
an empty or mismatched input, MSMError.noGPU when no device exists,
MSMError.gpuError otherwise).  Every compute call goes through the HIP
library; there is no CPU path.  Importing works without a GPU; creating an
engine does not.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzkmetal_amd.so")

SUCCESS, ERR_NO_GPU, ERR_INVALID_INPUT, ERR_GPU_ERROR, ERR_ALLOC_FAILED = 0, -1, -2, -3, -4

_lib = None


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("msms_per_s", ctypes.c_double), ("mean_batch", ctypes.c_double),
                ("hbm_peak_bytes", ctypes.c_uint64), ("hbm_in_use_bytes", ctypes.c_uint64)]


class MSMError(Exception):
    """Mirror of the reference's MSMError (MSMEngine.swift:8-15)."""

    def __init__(self, status, what=""):
        self.status = status
        names = {ERR_NO_GPU: "noGPU", ERR_INVALID_INPUT: "invalidInput",
                 ERR_GPU_ERROR: "gpuError", ERR_ALLOC_FAILED: "allocFailed"}
        super().__init__(f"{names.get(status, status)} {what}".strip())


def lib():
    """Load libzkmetal_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: build it with `python -m zkmetal_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, u8p, u32p = ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)
    i32, u32, u64 = ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64
    sig = {
        "zkmetal_msm_engine_create": (i32, [ctypes.POINTER(vp)]),
        "zkmetal_msm_engine_destroy": (None, [vp]),
        "zkmetal_bn254_msm": (i32, [vp, vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_auto": (i32, [vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_u8": (i32, [vp, vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_u16": (i32, [vp, vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_u32": (i32, [vp, vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_u8_auto": (i32, [vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_u16_auto": (i32, [vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_u32_auto": (i32, [vp, vp, u32, vp, vp, vp]),
        "zkmetal_bn254_msm_batch": (i32, [vp, vp, vp, vp, u32, vp]),
        "zkmetal_bn254_msm_batch_auto": (i32, [vp, vp, vp, u32, vp]),
        "zkmetal_set_shader_dir": (None, [u8p]),
        "zkmetal_gpu_available": (i32, []),
        "zkmetal_version": (ctypes.c_char_p, []),
        "zkmsm_srs_upload": (i32, [vp, vp, u32]),
        "zkmsm_srs_generate": (i32, [vp, u64, u32]),
        "zkmsm_srs_download": (i32, [vp, vp, u32]),
        "zkmsm_srs_size": (u32, [vp]),
        "zkmsm_msm_batch_resident": (i32, [vp, vp, vp, u32, vp]),
        "zkmsm_msm_batch_resident_device": (i32, [vp, vp, vp, u32, vp]),
        "zkmsm_msm_batch_resident_device_noalloc": (i32, [vp, vp, vp, u32, vp]),
        "zkmsm_msm_batch_device": (i32, [vp, vp, vp, vp, u32, vp]),
        "zkmsm_device_alloc": (i32, [u64, ctypes.POINTER(vp)]),
        "zkmsm_device_free": (i32, [vp]),
        "zkmsm_host_alloc": (i32, [u64, ctypes.POINTER(vp)]),
        "zkmsm_host_free": (i32, [vp]),
        "zkmsm_copy_to_device": (i32, [vp, vp, u64]),
        "zkmsm_copy_to_host": (i32, [vp, vp, u64]),
        "zkmsm_set_window_bits": (i32, [vp, u32]),
        "zkmsm_default_window_bits": (u32, [u64]),
        "zkmsm_last_timing": (i32, [vp, ctypes.POINTER(ctypes.c_double), i32]),
        "zkmsm_gen_scalars": (None, [u64, u64, i32, vp]),
        "zkmsm_test_fp": (i32, [vp, i32, vp, vp, vp, u32]),
        "zkmsm_test_xyzz": (i32, [vp, i32, vp, vp, vp, u32]),
        "zkmsm_test_glv": (i32, [vp, vp, vp, u32]),
        "zkmsm_set_workspace_limit": (i32, [vp, u64]),
        "zkmsm_set_tail_mode": (i32, [vp, i32]),
        "zkmsm_set_window_shift": (i32, [vp, i32]),
        "zkmsm_srs_window_shift": (u32, [vp]),
        "zkmsm_set_pipeline_parts": (i32, [vp, u32]),
        "zkmsm_workspace_bytes": (u64, [vp]),
        "zkmsm_probe_rates": (i32, [vp, ctypes.POINTER(ctypes.c_double)]),
        "zkmsm_probe_copy": (i32, [vp, u64, ctypes.POINTER(ctypes.c_double)]),
        "zkmsm_set_window_table": (i32, [vp, vp, u32]),
        "zkmsm_set_batch_window_table": (i32, [vp, vp, u32]),
        "zkmsm_set_schedule_table": (i32, [vp, i32, vp, vp, u32]),
        "zkmsm_set_graph_mode": (i32, [vp, i32]),
        "zkmsm_kzg_commit": (i32, [vp, vp, vp, u32, vp]),
        "zkmsm_srs_load_ptau": (i32, [vp, vp, u64, u32, ctypes.POINTER(u32)]),
        "zkmsm_srs_load_ptau_file": (i32, [vp, ctypes.c_char_p, u32, ctypes.POINTER(u32)]),
        "zkmsm_kzg_commit_device": (i32, [vp, vp, vp, u32, vp]),
        "zkmsm_graph_captures": (u64, [vp]),
        "zkmsm_server_create": (i32, [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "zkmsm_server_submit": (i32, [vp, vp, u32, ctypes.POINTER(u64)]),
        "zkmsm_server_submit_ref": (i32, [vp, vp, u32, ctypes.POINTER(u64)]),
        "zkmsm_server_wait": (i32, [vp, u64, vp, ctypes.POINTER(ctypes.c_double)]),
        "zkmsm_server_poll": (i32, [vp, u64, vp, ctypes.POINTER(i32)]),
        "zkmsm_server_stats": (i32, [vp, ctypes.POINTER(ServerStats)]),
        "zkmsm_server_destroy": (None, [vp]),
        "zkmsm_server_footprint": (u64, [u32]),
        "zkmsm_multi_msm": (i32, [vp, vp, u32, vp, u32, vp]),
        "zkmsm_set_srs_detect": (i32, [vp, i32]),
        "zkmsm_grumpkin_msm": (i32, [vp, vp, vp, u32, vp, vp, vp]),
        "zkmsm_grumpkin_msm_batch": (i32, [vp, vp, vp, vp, u32, vp]),
        "zkmsm_grumpkin_msm_batch_device": (i32, [vp, vp, vp, vp, u32, vp]),
        "zkmsm_grumpkin_points_generate": (i32, [vp, u64, u32, vp]),
        "zkmsm_bn254_g2_msm": (i32, [vp, vp, vp, u32, vp]),
        "zkmsm_bn254_g2_msm_batch": (i32, [vp, vp, vp, vp, u32, vp]),
        "zkmsm_bn254_g2_msm_batch_device": (i32, [vp, vp, vp, vp, u32, vp]),
        "zkmsm_bn254_g2_points_generate": (i32, [vp, u64, u32, vp]),
        "zkmsm_test_g2": (i32, [vp, i32, vp, vp, vp, u32]),
        "zkmsm_srs_detect_hits": (u64, [vp]),
        "zkmsm_inner_product_msm": (i32, [vp, vp, vp, u32, vp, u32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _buf(x):
    """bytes / bytearray / memoryview / numpy -> (ctypes pointer, keepalive),
    without copying the data (the library only reads inputs)."""
    if isinstance(x, np.ndarray):
        a = np.ascontiguousarray(x)
        return ctypes.c_void_p(a.ctypes.data), a
    a = np.frombuffer(x, dtype=np.uint8)
    if a.size == 0:
        return ctypes.c_void_p(0), a
    return ctypes.c_void_p(a.ctypes.data), a


def _check(rc, what=""):
    if rc != SUCCESS:
        raise MSMError(rc, what)


def _nbytes(x):
    """bytes held by a host buffer (bytes-like / numpy) or a DeviceBuffer"""
    if isinstance(x, DeviceBuffer):
        return x.nbytes
    if isinstance(x, np.ndarray):
        return x.nbytes
    return memoryview(x).nbytes


def _counts(counts):
    """counts as a contiguous uint32 array; negative or >= 2^32 is an error"""
    c = np.asarray(counts, dtype=np.int64).reshape(-1)
    if c.size and (c.min() < 0 or c.max() >= 1 << 32):
        raise MSMError(ERR_INVALID_INPUT, "counts out of range")
    return np.ascontiguousarray(c.astype(np.uint32))


def _need(buf, need, what):
    """The C ABI reads sum(counts) elements from each buffer: refuse a call
    whose buffer is shorter, before any device work is launched."""
    have = _nbytes(buf)
    if need > have:
        raise MSMError(ERR_INVALID_INPUT, f"{what}: buffer holds {have} bytes, the counts need {need}")


def gpu_available():
    return bool(lib().zkmetal_gpu_available())


def version():
    return lib().zkmetal_version().decode()


def gen_scalars(seed, n, dist=0):
    """n seeded scalars (32 B each) as a (n, 32) uint8 array."""
    out = np.zeros((n, 32), dtype=np.uint8)
    lib().zkmsm_gen_scalars(seed, n, dist, ctypes.c_void_p(out.ctypes.data))
    return out


def default_window_bits(n):
    return int(lib().zkmsm_default_window_bits(n))


class DeviceBuffer:
    """Device memory owned by Python (scalars kept resident in HBM)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        _check(lib().zkmsm_device_alloc(nbytes, ctypes.byref(self.ptr)), "device_alloc")
        self.nbytes = nbytes

    def upload(self, host):
        p, keep = _buf(host)
        _check(lib().zkmsm_copy_to_device(self.ptr, p, self.nbytes), "copy_to_device")
        del keep

    def download(self, nbytes=None):
        nbytes = self.nbytes if nbytes is None else nbytes
        out = np.zeros(nbytes, dtype=np.uint8)
        _check(lib().zkmsm_copy_to_host(ctypes.c_void_p(out.ctypes.data), self.ptr, nbytes),
               "copy_to_host")
        return out.tobytes()

    def free(self):
        if self.ptr:
            lib().zkmsm_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer:
    """Page-locked host memory; `.array` is a uint8 numpy view of it."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        _check(lib().zkmsm_host_alloc(nbytes, ctypes.byref(self.ptr)), "host_alloc")
        self.nbytes = nbytes
        raw = (ctypes.c_uint8 * nbytes).from_address(self.ptr.value)
        self.array = np.frombuffer(raw, dtype=np.uint8)

    def free(self):
        if self.ptr:
            self.array = None
            lib().zkmsm_host_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


TIMING_KEYS = ["decompose_ms", "sort_ms", "accumulate_ms", "merge_ms", "reduce_ms", "copyout_ms",
               "device_ms", "host_tail_ms", "wall_ms", "window_bits", "entries_per_lane",
               "buckets_per_group", "entries", "lanes", "valid_entries", "device_tail_ms", "parts"]


class MetalMSM:
    """BN254 G1 MSM engine (one HIP stream + device workspace)."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        _check(lib().zkmetal_msm_engine_create(ctypes.byref(self.handle)), "engine_create")

    def close(self):
        if self.handle:
            lib().zkmetal_msm_engine_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def msm(self, points, scalars, n=None):
        """points: n*64 B affine Montgomery; scalars: n*32 B standard form.
        Returns the 96-byte Jacobian (X|Y|Z), Montgomery, normalised."""
        n = _count(points, 64) if n is None else n
        if n == 0 or _count(scalars, 32) != n or _count(points, 64) != n:
            raise MSMError(ERR_INVALID_INPUT, "points/scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = (ctypes.c_uint8 * 96)()
        base = ctypes.addressof(out)
        _check(lib().zkmetal_bn254_msm(self.handle, pp, sp, n, base, base + 32, base + 64), "msm")
        return bytes(out)

    def msm_small(self, points, scalars, width):
        if width not in (1, 2, 4):
            raise MSMError(ERR_INVALID_INPUT, "msm_small: width is 1, 2 or 4 bytes")
        n = _count(points, 64)
        _need(scalars, n * width, "msm_small scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = (ctypes.c_uint8 * 96)()
        base = ctypes.addressof(out)
        f = {1: lib().zkmetal_bn254_msm_u8, 2: lib().zkmetal_bn254_msm_u16, 4: lib().zkmetal_bn254_msm_u32}[width]
        _check(f(self.handle, pp, sp, n, base, base + 32, base + 64), "msm_small")
        return bytes(out)

    def msm_batch(self, points, scalars, counts):
        """zkmetal_bn254_msm_batch: concatenated inputs -> list of 96-byte results."""
        counts = _counts(counts)
        total = int(counts.sum(dtype=np.uint64))
        _need(points, total * 64, "msm_batch points")
        _need(scalars, total * 32, "msm_batch scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        _check(lib().zkmetal_bn254_msm_batch(self.handle, pp, sp, ctypes.c_void_p(counts.ctypes.data),
                                             len(counts), ctypes.c_void_p(out.ctypes.data)), "msm_batch")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]

    def grumpkin_msm(self, points, scalars):
        """zkmsm_grumpkin_msm: Grumpkin points (x, y in Fr Montgomery form) and
        scalars (standard form) -> 96-byte Jacobian X|Y|Z (Fr Montgomery,
        normalised)."""
        n = _count(points, 64)
        if n == 0 or _count(scalars, 32) != n:
            raise MSMError(ERR_INVALID_INPUT, "points/scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = (ctypes.c_uint8 * 96)()
        base = ctypes.addressof(out)
        _check(lib().zkmsm_grumpkin_msm(self.handle, pp, sp, n, base, base + 32, base + 64),
               "grumpkin_msm")
        return bytes(out)

    def grumpkin_msm_batch(self, points, scalars, counts):
        counts = _counts(counts)
        total = int(counts.sum(dtype=np.uint64))
        _need(points, total * 64, "grumpkin_msm_batch points")
        _need(scalars, total * 32, "grumpkin_msm_batch scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        _check(lib().zkmsm_grumpkin_msm_batch(self.handle, pp, sp, ctypes.c_void_p(counts.ctypes.data),
                                              len(counts), ctypes.c_void_p(out.ctypes.data)),
               "grumpkin_msm_batch")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]

    def grumpkin_msm_batch_device(self, dpoints, dscalars, counts):
        """Grumpkin MSMs with bases and scalars in DeviceBuffers."""
        counts = _counts(counts)
        total = int(counts.sum(dtype=np.uint64))
        _need(dpoints, total * 64, "grumpkin_msm_batch_device points")
        _need(dscalars, total * 32, "grumpkin_msm_batch_device scalars")
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        _check(lib().zkmsm_grumpkin_msm_batch_device(self.handle, dpoints.ptr, dscalars.ptr,
                                                     ctypes.c_void_p(counts.ctypes.data), len(counts),
                                                     ctypes.c_void_p(out.ctypes.data)),
               "grumpkin_msm_batch_device")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]

    def grumpkin_points_generate(self, seed, n, dpoints):
        """n Grumpkin bases k_i * G (k_i = gen_scalars(seed, n, 0)) into a
        DeviceBuffer of at least n * 64 bytes."""
        if dpoints.nbytes < 64 * n:
            raise MSMError(ERR_INVALID_INPUT, "device buffer too small")
        _check(lib().zkmsm_grumpkin_points_generate(self.handle, seed, n, dpoints.ptr),
               "grumpkin_points_generate")

    def g2_msm(self, points, scalars):
        """zkmsm_bn254_g2_msm: G2 points (128 B: x.c0, x.c1, y.c0, y.c1, Fp
        Montgomery) and scalars (standard form) -> 192-byte Jacobian
        X | Y | Z (each c0 | c1), normalised."""
        n = _count(points, 128)
        if n == 0 or _count(scalars, 32) != n:
            raise MSMError(ERR_INVALID_INPUT, "points/scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = np.zeros(192, dtype=np.uint8)
        _check(lib().zkmsm_bn254_g2_msm(self.handle, pp, sp, n, ctypes.c_void_p(out.ctypes.data)), "g2_msm")
        return out.tobytes()

    def g2_msm_batch(self, points, scalars, counts):
        counts = _counts(counts)
        total = int(counts.sum(dtype=np.uint64))
        _need(points, total * 128, "g2_msm_batch points")
        _need(scalars, total * 32, "g2_msm_batch scalars")
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = np.zeros(len(counts) * 192, dtype=np.uint8)
        _check(lib().zkmsm_bn254_g2_msm_batch(self.handle, pp, sp, ctypes.c_void_p(counts.ctypes.data),
                                              len(counts), ctypes.c_void_p(out.ctypes.data)),
               "g2_msm_batch")
        return [out[i * 192:(i + 1) * 192].tobytes() for i in range(len(counts))]

    def g2_msm_batch_device(self, dpoints, dscalars, counts):
        counts = _counts(counts)
        total = int(counts.sum(dtype=np.uint64))
        _need(dpoints, total * 128, "g2_msm_batch_device points")
        _need(dscalars, total * 32, "g2_msm_batch_device scalars")
        out = np.zeros(len(counts) * 192, dtype=np.uint8)
        _check(lib().zkmsm_bn254_g2_msm_batch_device(self.handle, dpoints.ptr, dscalars.ptr,
                                                     ctypes.c_void_p(counts.ctypes.data), len(counts),
                                                     ctypes.c_void_p(out.ctypes.data)),
               "g2_msm_batch_device")
        return [out[i * 192:(i + 1) * 192].tobytes() for i in range(len(counts))]

    def g2_points_generate(self, seed, n, dpoints):
        """n G2 bases k_i * G2 (k_i = gen_scalars(seed, n, 0)) into a
        DeviceBuffer of at least n * 128 bytes."""
        if dpoints.nbytes < 128 * n:
            raise MSMError(ERR_INVALID_INPUT, "device buffer too small")
        _check(lib().zkmsm_bn254_g2_points_generate(self.handle, seed, n, dpoints.ptr),
               "g2_points_generate")

    def test_g2(self, op, p, q):
        n = _count(p, 256)
        out = np.zeros(n * 256, dtype=np.uint8)
        pp, k1 = _buf(p)
        qp, k2 = _buf(q)
        _check(lib().zkmsm_test_g2(self.handle, op, pp, qp, ctypes.c_void_p(out.ctypes.data), n), "test_g2")
        return out.tobytes()

    def set_srs_detect(self, on):
        """Serve host-point calls whose bases are prefixes of the resident SRS
        from the resident table (default on)."""
        _check(lib().zkmsm_set_srs_detect(self.handle, 1 if on else 0), "set_srs_detect")

    def srs_detect_hits(self):
        return int(lib().zkmsm_srs_detect_hits(self.handle))

    def multi_msm(self, points, scalar_sets):
        """zkmsm_multi_msm: k scalar sets (k x n x 32 B) over one base set of n
        points -> list of k 96-byte results."""
        n = _count(points, 64)
        sp, k2 = _buf(scalar_sets)
        total = _count(scalar_sets, 32)
        if n == 0 or total % n:
            raise MSMError(ERR_INVALID_INPUT, "multi_msm: scalar sets must be k x n")
        k = total // n
        pp, k1 = _buf(points)
        out = np.zeros(k * 96, dtype=np.uint8)
        _check(lib().zkmsm_multi_msm(self.handle, pp, n, sp, k, ctypes.c_void_p(out.ctypes.data)),
               "multi_msm")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(k)]

    def inner_product_msm(self, points, scalars, partitions):
        """zkmsm_inner_product_msm: one MSM per range [partitions[j], partitions[j+1])."""
        n = _count(points, 64)
        parts = np.ascontiguousarray(np.asarray(partitions, dtype=np.uint32))
        nparts = max(len(parts) - 1, 0)
        pp, k1 = _buf(points)
        sp, k2 = _buf(scalars)
        out = np.zeros(max(nparts, 1) * 96, dtype=np.uint8)
        _check(lib().zkmsm_inner_product_msm(self.handle, pp, sp, n, ctypes.c_void_p(parts.ctypes.data),
                                             nparts, ctypes.c_void_p(out.ctypes.data)),
               "inner_product_msm")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(nparts)]

    def msm_batch_device(self, dpoints, dscalars, counts):
        """Plain MSMs with bases and scalars in DeviceBuffers."""
        counts = _counts(counts)
        total = int(counts.sum(dtype=np.uint64))
        _need(dpoints, total * 64, "msm_batch_device points")
        _need(dscalars, total * 32, "msm_batch_device scalars")
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        _check(lib().zkmsm_msm_batch_device(self.handle, dpoints.ptr, dscalars.ptr,
                                            ctypes.c_void_p(counts.ctypes.data), len(counts),
                                            ctypes.c_void_p(out.ctypes.data)), "msm_batch_device")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]

    # ---- resident base table ----
    def srs_upload(self, points):
        n = _count(points, 64)
        pp, k = _buf(points)
        _check(lib().zkmsm_srs_upload(self.handle, pp, n), "srs_upload")

    def srs_generate(self, seed, n):
        _check(lib().zkmsm_srs_generate(self.handle, seed, n), "srs_generate")

    def srs_download(self, n):
        out = np.zeros(n * 64, dtype=np.uint8)
        _check(lib().zkmsm_srs_download(self.handle, ctypes.c_void_p(out.ctypes.data), n), "srs_download")
        return out

    def msm_batch_resident(self, scalars, counts):
        counts = _counts(counts)
        _need(scalars, int(counts.sum(dtype=np.uint64)) * 32, "msm_batch_resident scalars")
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        if isinstance(scalars, DeviceBuffer):
            rc = lib().zkmsm_msm_batch_resident_device(self.handle, scalars.ptr, ctypes.c_void_p(counts.ctypes.data),
                                                       len(counts), ctypes.c_void_p(out.ctypes.data))
        else:
            sp, k = _buf(scalars)
            rc = lib().zkmsm_msm_batch_resident(self.handle, sp, ctypes.c_void_p(counts.ctypes.data),
                                                len(counts), ctypes.c_void_p(out.ctypes.data))
        _check(rc, "msm_batch_resident")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]

    def srs_load_ptau(self, data=None, path=None, max_points=0):
        """Resident table from an SRS file (reference .ptau layout): bytes or a
        path; returns the number of bases loaded."""
        n = ctypes.c_uint32()
        if path is not None:
            rc = lib().zkmsm_srs_load_ptau_file(self.handle, os.fsencode(path), max_points, ctypes.byref(n))
        else:
            sp, k = _buf(data)
            rc = lib().zkmsm_srs_load_ptau(self.handle, sp, len(data), max_points, ctypes.byref(n))
        _check(rc, "srs_load_ptau")
        return n.value

    def kzg_commit(self, coeffs_mont, counts):
        """KZG commitments over the resident SRS (KZGEngine.batchCommit): each
        polynomial's coefficients are Fr elements in Montgomery form (32 B LE),
        concatenated; one 96-byte X|Y|Z result per polynomial."""
        counts = _counts(counts)
        _need(coeffs_mont, int(counts.sum(dtype=np.uint64)) * 32, "kzg_commit scalars")
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        if isinstance(coeffs_mont, DeviceBuffer):
            rc = lib().zkmsm_kzg_commit_device(self.handle, coeffs_mont.ptr, ctypes.c_void_p(counts.ctypes.data),
                                               len(counts), ctypes.c_void_p(out.ctypes.data))
        else:
            sp, k = _buf(coeffs_mont)
            rc = lib().zkmsm_kzg_commit(self.handle, sp, ctypes.c_void_p(counts.ctypes.data),
                                        len(counts), ctypes.c_void_p(out.ctypes.data))
        _check(rc, "kzg_commit")
        return [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]

    def set_window_table(self, table, batch=False):
        """table[ceil(log2 n)] = window width for that size class (0 = default);
        batch=True sets the table used for MSMs inside multi-MSM calls."""
        arr = np.ascontiguousarray(np.asarray(table, dtype=np.uint32))
        fn = lib().zkmsm_set_batch_window_table if batch else lib().zkmsm_set_window_table
        _check(fn(self.handle, ctypes.c_void_p(arr.ctypes.data), len(arr)), "set_window_table")

    def set_schedule_table(self, parts=None, reduce_per=None, batch=False):
        """Launch schedule by size class ceil(log2 entries) of a call:
        parts[k] pipeline parts (1..8), reduce_per[k] buckets per
        bucket-reduction lane (a power of two in 4..64); 0 (or None for a
        whole table) keeps the built-in rule.  batch=True sets the table for
        calls of several MSMs."""
        arrs = []
        for t in (parts, reduce_per):
            arrs.append(None if t is None else np.ascontiguousarray(np.asarray(t, dtype=np.uint32)))
        lens = [len(a) for a in arrs if a is not None]
        if len(set(lens)) > 1:
            raise MSMError(ERR_INVALID_INPUT, "set_schedule_table: tables of different lengths")
        ptrs = [None if a is None else ctypes.c_void_p(a.ctypes.data) for a in arrs]
        _check(lib().zkmsm_set_schedule_table(self.handle, int(bool(batch)), ptrs[0], ptrs[1],
                                              lens[0] if lens else 0), "set_schedule_table")

    GRAPH_OFF, GRAPH_ON, GRAPH_AUTO = 0, 1, 2

    def set_graph_mode(self, mode):
        """GRAPH_ON (or True): capture each call's device launch sequence in a
        hipGraph and replay it for later calls of the same shape; GRAPH_AUTO
        (the default): the same from the second call in a row of one shape;
        GRAPH_OFF (or False): every call eagerly."""
        _check(lib().zkmsm_set_graph_mode(self.handle, int(mode)), "set_graph_mode")

    TAIL_AUTO, TAIL_HOST, TAIL_DEVICE = 0, 1, 2

    def set_tail_mode(self, mode):
        """Where the window combine and normalisation run: TAIL_AUTO (device in
        graph mode and for calls of >= 16 MSMs), TAIL_HOST or TAIL_DEVICE."""
        _check(lib().zkmsm_set_tail_mode(self.handle, mode), "set_tail_mode")

    SHIFT_OFF, SHIFT_AUTO, SHIFT_ALL = 0, 1, 2

    def set_window_shift(self, mode):
        """Window-shifted resident tables: SHIFT_AUTO (default) or SHIFT_ALL
        make an SRS set afterwards also hold 2^(16 w) P_i for every window w,
        and MSMs on it share one bucket set across windows (auto: when the
        cost estimate favours it; all: always); SHIFT_OFF runs plain and
        makes the next SRS plain.  True / False mean auto / off.  Results do
        not change."""
        _check(lib().zkmsm_set_window_shift(self.handle, int(mode)), "set_window_shift")

    def srs_window_shift(self):
        """16 when the resident table holds window-shifted copies, else 0."""
        return int(lib().zkmsm_srs_window_shift(self.handle))

    def set_pipeline_parts(self, parts):
        """Parts a large call is split into for the sort / accumulation
        overlap (0 = by size, 1..8 = that many); results do not change."""
        _check(lib().zkmsm_set_pipeline_parts(self.handle, parts), "set_pipeline_parts")

    def set_workspace_limit(self, nbytes):
        """Cap the device memory this engine's buffers may grow to (0 = none);
        a call needing more raises MSMError(ERR_ALLOC_FAILED)."""
        _check(lib().zkmsm_set_workspace_limit(self.handle, nbytes), "set_workspace_limit")

    def workspace_bytes(self):
        return int(lib().zkmsm_workspace_bytes(self.handle))

    def graph_captures(self):
        return int(lib().zkmsm_graph_captures(self.handle))

    def set_window_bits(self, c):
        _check(lib().zkmsm_set_window_bits(self.handle, c), "set_window_bits")

    def last_timing(self):
        v = (ctypes.c_double * 17)()
        k = lib().zkmsm_last_timing(self.handle, v, 17)
        return {TIMING_KEYS[i]: v[i] for i in range(k)}

    def probe_rates(self):
        """Measured multiply throughputs (full occupancy) and single-lane latencies."""
        v = (ctypes.c_double * 6)()
        _check(lib().zkmsm_probe_rates(self.handle, v), "probe_rates")
        return {"mad64_per_s": v[0], "fips_step_per_s": v[1], "fpmul_per_s": v[2], "fpmul_chain_ns": v[3],
                "fq29mul_per_s": v[4], "fq29mul_chain_ns": v[5]}

    def probe_copy(self, nbytes):
        """ms of one device copy of nbytes (HBM counter calibration)."""
        ms = ctypes.c_double()
        _check(lib().zkmsm_probe_copy(self.handle, nbytes, ctypes.byref(ms)), "probe_copy")
        return ms.value

    # ---- primitive probes (GPU unit tests) ----
    def test_fp(self, op, a, b):
        n = _count(a, 32)
        ap, k1 = _buf(a)
        bp, k2 = _buf(b)
        out = np.zeros(n * 32, dtype=np.uint8)
        _check(lib().zkmsm_test_fp(self.handle, op, ap, bp, ctypes.c_void_p(out.ctypes.data), n), "test_fp")
        return out

    def test_xyzz(self, op, p, q):
        n = _count(p, 128)
        pp, k1 = _buf(p)
        qp, k2 = _buf(q)
        out = np.zeros(n * 128, dtype=np.uint8)
        _check(lib().zkmsm_test_xyzz(self.handle, op, pp, qp, ctypes.c_void_p(out.ctypes.data), n), "test_xyzz")
        return out

    def test_glv(self, scalars):
        n = _count(scalars, 32)
        sp, k = _buf(scalars)
        out = np.zeros(n * 48, dtype=np.uint8)
        _check(lib().zkmsm_test_glv(self.handle, sp, ctypes.c_void_p(out.ctypes.data), n), "test_glv")
        return out


def _count(x, size):
    nb = x.nbytes if isinstance(x, np.ndarray) else len(x)
    return nb // size


class Server:
    """Continuous-batching request server over an engine's resident SRS
    (zkmsm_server_*).  submit() returns a ticket; wait()/poll() return the
    96-byte result."""

    def __init__(self, engine, slots=2, max_batch=64, hbm_budget_bytes=0):
        cfg = ServerConfig(slots, max_batch, hbm_budget_bytes)
        self.handle = ctypes.c_void_p()
        self._engine = engine  # keeps the SRS owner alive
        _check(lib().zkmsm_server_create(engine.handle, ctypes.byref(cfg), ctypes.byref(self.handle)),
               "server_create")

    def submit(self, scalars, copy=True):
        """copy=False reads the scalars in place: pass a contiguous numpy array
        and keep it alive and unchanged until the request is done."""
        sp, keep = _buf(scalars)
        t = ctypes.c_uint64()
        if copy:
            rc = lib().zkmsm_server_submit(self.handle, sp, _count(scalars, 32), ctypes.byref(t))
        else:
            if not (isinstance(scalars, np.ndarray) and scalars.flags["C_CONTIGUOUS"]):
                raise MSMError(ERR_INVALID_INPUT, "submit(copy=False) needs a contiguous numpy array")
            rc = lib().zkmsm_server_submit_ref(self.handle, sp, _count(scalars, 32), ctypes.byref(t))
        _check(rc, "server_submit")
        return t.value

    def wait(self, ticket):
        out = (ctypes.c_uint8 * 96)()
        lat = ctypes.c_double()
        _check(lib().zkmsm_server_wait(self.handle, ticket, out, ctypes.byref(lat)), "server_wait")
        return bytes(out), lat.value

    def poll(self, ticket):
        out = (ctypes.c_uint8 * 96)()
        done = ctypes.c_int32()
        _check(lib().zkmsm_server_poll(self.handle, ticket, out, ctypes.byref(done)), "server_poll")
        return bytes(out) if done.value else None

    def stats(self):
        s = ServerStats()
        _check(lib().zkmsm_server_stats(self.handle, ctypes.byref(s)), "server_stats")
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().zkmsm_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_footprint(n):
    return int(lib().zkmsm_server_footprint(n))
