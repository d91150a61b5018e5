"""ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Wraps the C restatement of the reference CPU Pippenger
(Sources/NeonFieldOps/bn254_msm.c:1145-1187) plus the helper routines the
tests use.  `curve="grumpkin"` selects oracle/liboracle_grumpkin.so, the same
source built over Grumpkin's field (grumpkin_ops.c:528-560); its field and
point helpers then work in BN254 Fr and it has no GLV routines.  The g2_*
functions wrap the BN254 G2 Pippenger of the same source (liboracle.so only).  Build with
`make -C oracle`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
_FILES = {"bn254": "liboracle.so", "grumpkin": "liboracle_grumpkin.so"}


def lib(curve="bn254"):
    if curve not in _LIBS:
        path = os.path.join(_HERE, _FILES[curve])
        if not os.path.exists(path):
            raise RuntimeError(f"oracle/{_FILES[curve]} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_pippenger_msm.argtypes = [u64p, u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u64p]
        L.oracle_jacobian_to_affine.argtypes = [u64p, u64p]
        L.oracle_scalar_mul.argtypes = [u64p, u32p, u64p]
        L.oracle_scalar_dot.argtypes = [u32p, u32p, ctypes.c_int, u32p]
        if curve == "bn254":
            L.oracle_glv_decompose.argtypes = [u32p, u32p, u32p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_signed_digits.argtypes = [u32p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int)]
        L.oracle_window_bits.argtypes = [ctypes.c_int]
        for name in ("oracle_fp_mul", "oracle_fp_add", "oracle_fp_sub"):
            getattr(L, name).argtypes = [u64p, u64p, u64p]
        for name in ("oracle_fp_inv", "oracle_fp_sqr", "oracle_to_mont", "oracle_from_mont"):
            getattr(L, name).argtypes = [u64p, u64p]
        L.oracle_jac_add.argtypes = [u64p, u64p, u64p]
        L.oracle_jac_dbl.argtypes = [u64p, u64p]
        L.oracle_jac_add_affine.argtypes = [u64p, u64p, u64p]
        if curve == "bn254":
            L.oracle_g2_pippenger_msm.argtypes = [u64p, u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, u64p]
            L.oracle_g2_jacobian_to_affine.argtypes = [u64p, u64p]
            L.oracle_g2_jac_add.argtypes = [u64p, u64p, u64p]
            L.oracle_g2_jac_dbl.argtypes = [u64p, u64p]
            L.oracle_g2_jac_add_affine.argtypes = [u64p, u64p, u64p]
        _LIBS[curve] = L
    return _LIBS[curve]


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def pippenger(points_u8, scalars_u8, n, window_bits=0, threads=0, curve="bn254"):
    """points_u8: n*64 bytes, scalars_u8: n*32 bytes -> (X, Y, Z) 96 bytes."""
    pts = np.frombuffer(bytes(points_u8), dtype=np.uint64) if not isinstance(points_u8, np.ndarray) else points_u8.view(np.uint64)
    sc = np.frombuffer(bytes(scalars_u8), dtype=np.uint32) if not isinstance(scalars_u8, np.ndarray) else scalars_u8.view(np.uint32)
    pts = np.ascontiguousarray(pts)
    sc = np.ascontiguousarray(sc)
    out = np.zeros(12, dtype=np.uint64)
    lib(curve).oracle_pippenger_msm(_p(pts, ctypes.c_uint64), _p(sc, ctypes.c_uint32), int(n), int(window_bits), int(threads), _p(out, ctypes.c_uint64))
    return out.tobytes()


def g2_pippenger(points_u8, scalars_u8, n, window_bits=0, threads=0):
    """BN254 G2: points_u8 n*128 bytes (x.c0, x.c1, y.c0, y.c1, Fp Montgomery),
    scalars_u8 n*32 bytes -> Jacobian (X, Y, Z), each c0 | c1, 192 bytes."""
    pts = np.ascontiguousarray(np.frombuffer(bytes(points_u8), dtype=np.uint64)
                               if not isinstance(points_u8, np.ndarray) else points_u8.view(np.uint64))
    sc = np.ascontiguousarray(np.frombuffer(bytes(scalars_u8), dtype=np.uint32)
                              if not isinstance(scalars_u8, np.ndarray) else scalars_u8.view(np.uint32))
    out = np.zeros(24, dtype=np.uint64)
    lib().oracle_g2_pippenger_msm(_p(pts, ctypes.c_uint64), _p(sc, ctypes.c_uint32), int(n),
                                  int(window_bits), int(threads), _p(out, ctypes.c_uint64))
    return out.tobytes()


def g2_jacobian_to_affine_bytes(jac192):
    j = np.frombuffer(bytes(jac192), dtype=np.uint64).copy()
    out = np.zeros(16, dtype=np.uint64)
    lib().oracle_g2_jacobian_to_affine(_p(j, ctypes.c_uint64), _p(out, ctypes.c_uint64))
    return out.tobytes()


def g2_point_op(name, *args):
    """oracle_g2_jac_add / _dbl / _add_affine on byte strings (192-byte
    Jacobian, 128-byte affine) -> 192-byte Jacobian."""
    ins = [np.frombuffer(bytes(a), dtype=np.uint64).copy() for a in args]
    out = np.zeros(24, dtype=np.uint64)
    getattr(lib(), name)(*[_p(x, ctypes.c_uint64) for x in ins], _p(out, ctypes.c_uint64))
    return out.tobytes()


def jacobian_to_affine_bytes(jac96, curve="bn254"):
    j = np.frombuffer(bytes(jac96), dtype=np.uint64).copy()
    out = np.zeros(8, dtype=np.uint64)
    lib(curve).oracle_jacobian_to_affine(_p(j, ctypes.c_uint64), _p(out, ctypes.c_uint64))
    return out.tobytes()


def scalar_dot(scalars_u8, ks_u8, n, curve="bn254"):
    s = np.frombuffer(bytes(scalars_u8), dtype=np.uint32).copy()
    k = np.frombuffer(bytes(ks_u8), dtype=np.uint32).copy()
    out = np.zeros(8, dtype=np.uint32)
    lib(curve).oracle_scalar_dot(_p(s, ctypes.c_uint32), _p(k, ctypes.c_uint32), int(n), _p(out, ctypes.c_uint32))
    return int.from_bytes(out.tobytes(), "little")


def scalar_mul_affine(aff64, k, curve="bn254"):
    a = np.frombuffer(bytes(aff64), dtype=np.uint64).copy()
    kk = np.frombuffer(int(k).to_bytes(32, "little"), dtype=np.uint32).copy()
    out = np.zeros(12, dtype=np.uint64)
    lib(curve).oracle_scalar_mul(_p(a, ctypes.c_uint64), _p(kk, ctypes.c_uint32), _p(out, ctypes.c_uint64))
    return out.tobytes()


def glv_decompose(k):
    kk = np.frombuffer(int(k).to_bytes(32, "little"), dtype=np.uint32).copy()
    k1 = np.zeros(8, dtype=np.uint32)
    k2 = np.zeros(8, dtype=np.uint32)
    n1 = ctypes.c_int()
    n2 = ctypes.c_int()
    lib().oracle_glv_decompose(_p(kk, ctypes.c_uint32), _p(k1, ctypes.c_uint32), _p(k2, ctypes.c_uint32), ctypes.byref(n1), ctypes.byref(n2))
    return (int.from_bytes(k1.tobytes(), "little"), int.from_bytes(k2.tobytes(), "little"), bool(n1.value), bool(n2.value))


def signed_digits(k, c, nwin):
    kk = np.frombuffer(int(k).to_bytes(32, "little"), dtype=np.uint32).copy()
    d = np.zeros(nwin, dtype=np.int32)
    carry = ctypes.c_int()
    lib().oracle_signed_digits(_p(kk, ctypes.c_uint32), int(c), int(nwin), _p(d, ctypes.c_int32), ctypes.byref(carry))
    return [int(x) for x in d], int(carry.value)


def fp_op(name, *args, curve="bn254"):
    ins = [np.frombuffer(int(a).to_bytes(32, "little"), dtype=np.uint64).copy() for a in args]
    out = np.zeros(4, dtype=np.uint64)
    getattr(lib(curve), name)(*[_p(x, ctypes.c_uint64) for x in ins], _p(out, ctypes.c_uint64))
    return int.from_bytes(out.tobytes(), "little")
