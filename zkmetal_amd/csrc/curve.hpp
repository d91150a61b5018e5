// BN254 G1 (y^2 = x^3 + 3) point arithmetic in XYZZ coordinates:
// (X, Y, ZZ, ZZZ) represents the affine point (X/ZZ, Y/ZZZ); the identity is
// ZZ == 0.  Formulas are the standard short-Weierstrass XYZZ ones
// (madd-2008-s 8M+2S, add-2008-s 12M+2S, dbl-2008-s-1 6M+3S with a = 0).
//
// Exception handling follows the reference CPU oracle, not its GPU kernels
// (Sources/NeonFieldOps/bn254_msm.c:421-533 vs
// Sources/Shaders/geometry/bn254_curve.metal:83-99): the affine point (0,0)
// is the identity, P + P doubles, P + (-P) is the identity.
#pragma once
#include "field.hpp"

namespace zk {

struct Aff {
    Fp x, y;
};

struct Xyzz {
    Fp X, Y, ZZ, ZZZ;
};

__device__ __forceinline__ Xyzz xyzz_identity() {
    Xyzz r;
    r.X = fp_zero();
    r.Y = fp_zero();
    r.ZZ = fp_zero();
    r.ZZZ = fp_zero();
    return r;
}

__device__ __forceinline__ bool xyzz_is_id(const Xyzz& p) { return fp_is_zero(p.ZZ); }

__device__ __forceinline__ bool aff_is_id(const Aff& q) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) x |= q.x.v[i] | q.y.v[i];
    return x == 0;
}

__device__ __forceinline__ Xyzz xyzz_select(bool c, const Xyzz& a, const Xyzz& b) {
    Xyzz r;
    r.X = fp_select(c, a.X, b.X);
    r.Y = fp_select(c, a.Y, b.Y);
    r.ZZ = fp_select(c, a.ZZ, b.ZZ);
    r.ZZZ = fp_select(c, a.ZZZ, b.ZZZ);
    return r;
}

// 2*(x, y) for an affine, non-identity point (mdbl-2008-s-1)
__device__ __forceinline__ Xyzz xyzz_dbl_aff(const Fp& x, const Fp& y) {
    Fp U = fp_dbl(y);
    Fp V = fp_sqr(U);
    Fp W = fp_mul(U, V);
    Fp S = fp_mul(x, V);
    Fp xx = fp_sqr(x);
    Fp M = fp_add(fp_dbl(xx), xx);
    Xyzz r;
    r.X = fp_sub(fp_sqr(M), fp_dbl(S));
    r.Y = fp_sub(fp_mul(M, fp_sub(S, r.X)), fp_mul(W, y));
    r.ZZ = V;
    r.ZZZ = W;
    return r;
}

// 2*P (dbl-2008-s-1); identity stays identity (ZZ3 = V*ZZ1 = 0)
__device__ __forceinline__ Xyzz xyzz_dbl(const Xyzz& p) {
    Fp U = fp_dbl(p.Y);
    Fp V = fp_mul(U, U);
    Fp W = fp_mul(U, V);
    Fp S = fp_mul(p.X, V);
    Fp xx = fp_mul(p.X, p.X);
    Fp M = fp_add(fp_dbl(xx), xx);
    Xyzz r;
    r.X = fp_sub(fp_mul(M, M), fp_dbl(S));
    r.Y = fp_sub(fp_mul(M, fp_sub(S, r.X)), fp_mul(W, p.Y));
    r.ZZ = fp_mul(V, p.ZZ);
    r.ZZZ = fp_mul(W, p.ZZZ);
    return r;
}

// acc += (x2, y2).  q_id: the affine point is (0,0), i.e. the identity.
__device__ __forceinline__ void xyzz_madd(Xyzz& acc, const Fp& x2, const Fp& y2, bool q_id) {
    const bool p_id = xyzz_is_id(acc);
    Fp U2 = fp_mul(x2, acc.ZZ);
    Fp S2 = fp_mul(y2, acc.ZZZ);
    Fp P = fp_sub(U2, acc.X);
    Fp R = fp_sub(S2, acc.Y);
    const bool dbl = fp_is_zero(P) && fp_is_zero(R) && !p_id && !q_id;
    Fp PP = fp_sqr(P);
    Fp PPP = fp_mul(P, PP);
    Fp Q = fp_mul(acc.X, PP);
    Xyzz r;
    r.X = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
    r.Y = fp_sub(fp_mul(R, fp_sub(Q, r.X)), fp_mul(acc.Y, PPP));
    r.ZZ = fp_mul(acc.ZZ, PP);
    r.ZZZ = fp_mul(acc.ZZZ, PPP);
    if (__builtin_expect(dbl, 0)) r = xyzz_dbl_aff(x2, y2);
    // P == -Q falls out of the formulas: P = 0 gives ZZ3 = 0, the identity.
    Xyzz q;
    q.X = x2;
    q.Y = y2;
    q.ZZ = fp_one();
    q.ZZZ = q.ZZ;
    r = xyzz_select(p_id, q, r);
    acc = xyzz_select(q_id, acc, r);
}

// acc += q (both XYZZ)
__device__ __forceinline__ void xyzz_add(Xyzz& acc, const Xyzz& q) {
    const bool p_id = xyzz_is_id(acc);
    const bool q_id = xyzz_is_id(q);
    Fp U1 = fp_mul(acc.X, q.ZZ);
    Fp U2 = fp_mul(q.X, acc.ZZ);
    Fp S1 = fp_mul(acc.Y, q.ZZZ);
    Fp S2 = fp_mul(q.Y, acc.ZZZ);
    Fp P = fp_sub(U2, U1);
    Fp R = fp_sub(S2, S1);
    const bool dbl = fp_is_zero(P) && fp_is_zero(R) && !p_id && !q_id;
    Fp PP = fp_mul(P, P);
    Fp PPP = fp_mul(P, PP);
    Fp Q = fp_mul(U1, PP);
    Xyzz r;
    r.X = fp_sub(fp_sub(fp_mul(R, R), PPP), fp_dbl(Q));
    r.Y = fp_sub(fp_mul(R, fp_sub(Q, r.X)), fp_mul(S1, PPP));
    r.ZZ = fp_mul(fp_mul(acc.ZZ, q.ZZ), PP);
    r.ZZZ = fp_mul(fp_mul(acc.ZZZ, q.ZZZ), PPP);
    if (__builtin_expect(dbl, 0)) r = xyzz_dbl(acc);
    r = xyzz_select(p_id, q, r);
    acc = xyzz_select(q_id, acc, r);
}

// Global-memory layouts: an affine point is 64 B (x then y), an XYZZ point
// 128 B (X, Y, ZZ, ZZZ); both are loaded/stored as 16-byte vectors.
__device__ __forceinline__ Fp load_fp(const uint4* p) {
    uint4 a = p[0], b = p[1];
    Fp r;
    r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
    r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
    return r;
}

__device__ __forceinline__ void store_fp(uint4* p, const Fp& f) {
    p[0] = make_uint4(f.v[0], f.v[1], f.v[2], f.v[3]);
    p[1] = make_uint4(f.v[4], f.v[5], f.v[6], f.v[7]);
}

__device__ __forceinline__ Aff load_aff(const uint4* base, uint64_t idx) {
    const uint4* p = base + idx * 4;
    Aff a;
    a.x = load_fp(p);
    a.y = load_fp(p + 2);
    return a;
}

__device__ __forceinline__ void store_aff(uint4* base, uint64_t idx, const Aff& a) {
    uint4* p = base + idx * 4;
    store_fp(p, a.x);
    store_fp(p + 2, a.y);
}

__device__ __forceinline__ Xyzz load_xyzz(const uint4* base, uint64_t idx) {
    const uint4* p = base + idx * 8;
    Xyzz r;
    r.X = load_fp(p);
    r.Y = load_fp(p + 2);
    r.ZZ = load_fp(p + 4);
    r.ZZZ = load_fp(p + 6);
    return r;
}

__device__ __forceinline__ void store_xyzz(uint4* base, uint64_t idx, const Xyzz& a) {
    uint4* p = base + idx * 8;
    store_fp(p, a.X);
    store_fp(p + 2, a.Y);
    store_fp(p + 4, a.ZZ);
    store_fp(p + 6, a.ZZZ);
}

}  // namespace zk
