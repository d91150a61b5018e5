// BN254 G1 XYZZ arithmetic over the 29-bit-limb field (field29.hpp): the
// same formulas and exception rules as curve.hpp (madd-2008-s, add-2008-s,
// dbl-2008-s-1; (0,0) affine is the identity, P + P doubles, P + (-P) is the
// identity), with lazy reduction.  Bounds, in multiples of p:
//   products < ab/2^261 + p: < 1.5p for inputs < 8p, < 1.6p for inputs up to
//   8.5p;  stored X < 6.5p (4p outside the accumulation loop), Y < 4p;
//   stored ZZ, ZZZ < 2p;  affine table coordinates < p (canonical).
// Stored points (buckets, merge records, reduction partials) keep their
// normalised 29-bit limbs, 144 B per point; base-table entries likewise.
#pragma once
#include "curve.hpp"
#include "field29.hpp"
#include "plan.hpp"

namespace zk {

struct QAff {
    Fq x, y;
};

struct QXyzz {
    Fq X, Y, ZZ, ZZZ;
};

__device__ __forceinline__ Fq fq_zero() {
    Fq r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = 0;
    return r;
}

__device__ __forceinline__ Fq fq_select(bool c, const Fq& a, const Fq& b) {
    Fq r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = c ? a.v[i] : b.v[i];
    return r;
}

__device__ __forceinline__ QXyzz qxyzz_identity() {
    QXyzz r;
    r.X = fq_zero();
    r.Y = fq_zero();
    r.ZZ = fq_zero();
    r.ZZZ = fq_zero();
    return r;
}

__device__ __forceinline__ bool qaff_is_id(const QAff& q) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) x |= q.x.v[i] | q.y.v[i];
    return x == 0;
}

__device__ __forceinline__ QXyzz qxyzz_select(bool c, const QXyzz& a, const QXyzz& b) {
    QXyzz r;
    r.X = fq_select(c, a.X, b.X);
    r.Y = fq_select(c, a.Y, b.Y);
    r.ZZ = fq_select(c, a.ZZ, b.ZZ);
    r.ZZZ = fq_select(c, a.ZZZ, b.ZZZ);
    return r;
}

// ---- global memory.  Base-table entries: one 128-byte line holding the
// limbs of x, y and p - y (words 0-8, 9-17, 18-26; 27-31 zero), canonical,
// so a gather needs no re-slicing and a negative digit just takes p - y.
// XYZZ points: 256-bit integers (< 4p < 2^256), 128 B, as in curve.hpp.
// Table entry, 128 B: words 0-7 x[0..7], 8-15 y[0..7], 16-23 (p-y)[0..7],
// 24-25 (x[8], y[8]), 26-27 (x[8], (p-y)[8]), 28-31 zero.  A load for either
// sign is two 32-byte pieces and one 8-byte piece of the same 128-byte line.
__device__ __forceinline__ QAff load_qtab(const uint4* base, uint64_t idx, bool neg) {
    static_assert(kTableEntryBytes == 128, "entry layout");
    const uint4* p = base + idx * 8;
    const uint4* py = p + (neg ? 4 : 2);
    const uint4 x0 = p[0], x1 = p[1], y0 = py[0], y1 = py[1];
    const uint2 h = reinterpret_cast<const uint2*>(p)[neg ? 13 : 12];
    QAff a;
    a.x.v[0] = x0.x; a.x.v[1] = x0.y; a.x.v[2] = x0.z; a.x.v[3] = x0.w;
    a.x.v[4] = x1.x; a.x.v[5] = x1.y; a.x.v[6] = x1.z; a.x.v[7] = x1.w;
    a.y.v[0] = y0.x; a.y.v[1] = y0.y; a.y.v[2] = y0.z; a.y.v[3] = y0.w;
    a.y.v[4] = y1.x; a.y.v[5] = y1.y; a.y.v[6] = y1.z; a.y.v[7] = y1.w;
    a.x.v[8] = h.x;
    a.y.v[8] = h.y;
    return a;
}

// Buckets, merge records and reduction partials: the 36 limbs as they are
// (X, Y, ZZ, ZZZ, 9 words each; kBucketBytes = 144 per point), so a store
// needs no repacking into 256-bit words.  All-zero bytes are the identity.
__device__ __forceinline__ QXyzz load_qxyzz(const uint4* base, uint64_t idx) {
    static_assert(kBucketBytes == 144, "bucket layout");
    const uint4* p = base + idx * 9;
    uint32_t w[36];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint4 v = p[i];
        w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
    }
    QXyzz r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        r.X.v[i] = w[i];
        r.Y.v[i] = w[9 + i];
        r.ZZ.v[i] = w[18 + i];
        r.ZZZ.v[i] = w[27 + i];
    }
    return r;
}

__device__ __forceinline__ void store_qxyzz(uint4* base, uint64_t idx, const QXyzz& a) {
    uint32_t w[36];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        w[i] = a.X.v[i];
        w[9 + i] = a.Y.v[i];
        w[18 + i] = a.ZZ.v[i];
        w[27 + i] = a.ZZZ.v[i];
    }
    uint4* p = base + idx * 9;
#pragma unroll
    for (int i = 0; i < 9; i++) p[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// The 29-bit field arithmetic over the field F (field29.hpp) and the XYZZ
// curve arithmetic on y^2 = x^3 + b over it (b enters no formula here).
template <class F>
struct Arith {
    // limb i of K*p, normalised (folded at compile time inside unrolled loops)
    template <int K>
    static __device__ __forceinline__ uint32_t q_kp(int i) {
        uint64_t c = 0;
        uint32_t r = 0;
#pragma unroll
        for (int j = 0; j < 9; j++) {
            const uint64_t s = (uint64_t)K * F::p(j) + c;
            if (j == i) r = (uint32_t)s & ZQ_MASK;
            c = s >> 29;
        }
        return r;
    }

    // K*p with B * 2^29 borrowed into every limb below the top one: limb 0 is
    // kp_0 + B 2^29, limbs 1..7 kp_i + B 2^29 - B, limb 8 kp_8 - B (same value).
    // For a c whose limbs are sums of at most B normalised limbs (c_i <=
    // B (2^29 - 1)) and c_8 <= kp_8 - B, every limb of (this - c) is >= 0, so
    // K*p - c needs no borrow propagation.
    template <int K, int B = 1>
    static __device__ __forceinline__ uint32_t q_kpr(int i) {
        const uint32_t c = q_kp<K>(i);
        return i == 0 ? c + (uint32_t)B * (1u << 29)
                      : (i < 8 ? c + (uint32_t)B * (1u << 29) - (uint32_t)B : c - (uint32_t)B);
    }

    // Q - c + K*p limb by limb with no carry propagation (limbs < 2^29 + B 2^29,
    // value < Q + K*p): a multiplicand for fq_mul_sub / fq_mul whose other operand
    // has normalised limbs.  Q normalised; c as for q_kpr<K, B>.
    template <int K, int B = 1>
    static __device__ __forceinline__ Fq fq_sub_lazy(const Fq& q, const Fq& c) {
        Fq r;
#pragma unroll
        for (int i = 0; i < 9; i++) r.v[i] = q.v[i] + q_kpr<K, B>(i) - c.v[i];
        return r;
    }

    // a * b / 2^261 - c + K*p mod p (K > 0), or a * b / 2^261 (K == 0):
    // result < ab/2^261 + p + K*p with normalised limbs.  The limbs of K*p - c
    // (q_kpr<K, B>) enter the high columns of the product, where they cost one
    // subtraction each instead of a separate normalised subtraction.
    // a, b: limbs < 2^30, or a normalised and b limbs < 1.5 * 2^30 (column sums
    // then stay below 2^63.3); values with ab < p * 2^261 / 2 (e.g. both < 8p);
    // c (K > 0): as for q_kpr<K, B>.
    template <int K, int B = 1>
    static __device__ __forceinline__ Fq fq_mul_sub(const Fq& a, const Fq& b, const Fq& c) {
        uint32_t m[9];
        Fq t;
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int j = 0; j < k; j++) {
                qmac(acc, a.v[j], b.v[k - j]);
                qmac(acc, m[j], F::p(k - j));
            }
            qmac(acc, a.v[k], b.v[0]);
            m[k] = ((uint32_t)acc * F::kPinv) & ZQ_MASK;
            qmac(acc, m[k], F::p(0));
            acc >>= 29;
            carry_fence(acc);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
            if (K > 0) acc += q_kpr<K, B>(k - 9) - c.v[k - 9];
#pragma unroll
            for (int j = k - 8; j < 9; j++) {
                qmac(acc, a.v[j], b.v[k - j]);
                qmac(acc, m[j], F::p(k - j));
            }
            t.v[k - 9] = (uint32_t)acc & ZQ_MASK;
            acc >>= 29;
            carry_fence(acc);
        }
        t.v[8] = (uint32_t)acc + (K > 0 ? q_kpr<K, B>(8) - c.v[8] : 0u);
        return t;
    }

    // a * b / 2^261 mod p, result < ab/2^261 + p (normalised limbs).
    // a, b: limbs < 2^30, values with ab < p * 2^261 / 2 (e.g. both < 8p).
    static __device__ __forceinline__ Fq fq_mul(const Fq& a, const Fq& b) { return fq_mul_sub<0>(a, b, a); }

    // a^2 / 2^261 mod p: the symmetric products a_i a_j (i != j) once, from the
    // doubled limb 2 a_i (< 2^30), plus the squares: 45 + 81 multiply-adds.
    // Column bound: 4 doubled products < 2^59, one square < 2^58, 9 m*p products
    // < 2^58 and the carry: < 18 * 2^58 < 2^63.  a: normalised limbs, value < 8p.
    // fq_sqr_sub<K, B>(a, c) = a^2 / 2^261 - c + K*p, as fq_mul_sub.
    template <int K, int B = 1>
    static __device__ __forceinline__ Fq fq_sqr_sub(const Fq& a, const Fq& c) {
        uint32_t a2[9], m[9];
#pragma unroll
        for (int i = 0; i < 9; i++) a2[i] = a.v[i] << 1;
        Fq t;
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int i = 0; 2 * i < k; i++) qmac(acc, a2[i], a.v[k - i]);
            if ((k & 1) == 0) qmac(acc, a.v[k / 2], a.v[k / 2]);
#pragma unroll
            for (int j = 0; j < k; j++) qmac(acc, m[j], F::p(k - j));
            m[k] = ((uint32_t)acc * F::kPinv) & ZQ_MASK;
            qmac(acc, m[k], F::p(0));
            acc >>= 29;
            carry_fence(acc);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
            if (K > 0) acc += q_kpr<K, B>(k - 9) - c.v[k - 9];
#pragma unroll
            for (int i = k - 8; 2 * i < k; i++) qmac(acc, a2[i], a.v[k - i]);
            if ((k & 1) == 0) qmac(acc, a.v[k / 2], a.v[k / 2]);
#pragma unroll
            for (int j = k - 8; j < 9; j++) qmac(acc, m[j], F::p(k - j));
            t.v[k - 9] = (uint32_t)acc & ZQ_MASK;
            acc >>= 29;
            carry_fence(acc);
        }
        t.v[8] = (uint32_t)acc + (K > 0 ? q_kpr<K, B>(8) - c.v[8] : 0u);
        return t;
    }

    static __device__ __forceinline__ Fq fq_sqr(const Fq& a) { return fq_sqr_sub<0>(a, a); }

    // Two independent products computed column by column side by side (two
    // multiply-add chains per column, so each hides the other's latency):
    // r1 = fq_mul_sub<K1, B1>(a1, b1, c1), r2 = fq_mul_sub<K2, B2>(a2, b2, c2),
    // same operand conditions.
    template <int K1, int B1, int K2, int B2>
    static __device__ __forceinline__ void fq_mul_sub_x2(const Fq& a1, const Fq& b1, const Fq& c1, Fq& r1,
                                                         const Fq& a2, const Fq& b2, const Fq& c2, Fq& r2) {
        uint32_t m1[9], m2[9];
        uint64_t s1 = 0, s2 = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int j = 0; j < k; j++) {
                qmac(s1, a1.v[j], b1.v[k - j]);
                qmac(s2, a2.v[j], b2.v[k - j]);
                qmac(s1, m1[j], F::p(k - j));
                qmac(s2, m2[j], F::p(k - j));
            }
            qmac(s1, a1.v[k], b1.v[0]);
            qmac(s2, a2.v[k], b2.v[0]);
            m1[k] = ((uint32_t)s1 * F::kPinv) & ZQ_MASK;
            m2[k] = ((uint32_t)s2 * F::kPinv) & ZQ_MASK;
            qmac(s1, m1[k], F::p(0));
            qmac(s2, m2[k], F::p(0));
            s1 >>= 29;
            s2 >>= 29;
            carry_fence(s1);
            carry_fence(s2);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
            if (K1 > 0) s1 += q_kpr<K1, B1>(k - 9) - c1.v[k - 9];
            if (K2 > 0) s2 += q_kpr<K2, B2>(k - 9) - c2.v[k - 9];
#pragma unroll
            for (int j = k - 8; j < 9; j++) {
                qmac(s1, a1.v[j], b1.v[k - j]);
                qmac(s2, a2.v[j], b2.v[k - j]);
                qmac(s1, m1[j], F::p(k - j));
                qmac(s2, m2[j], F::p(k - j));
            }
            r1.v[k - 9] = (uint32_t)s1 & ZQ_MASK;
            r2.v[k - 9] = (uint32_t)s2 & ZQ_MASK;
            s1 >>= 29;
            s2 >>= 29;
            carry_fence(s1);
            carry_fence(s2);
        }
        r1.v[8] = (uint32_t)s1 + (K1 > 0 ? q_kpr<K1, B1>(8) - c1.v[8] : 0u);
        r2.v[8] = (uint32_t)s2 + (K2 > 0 ? q_kpr<K2, B2>(8) - c2.v[8] : 0u);
    }

    // Three independent products a_i b_i / 2^261 side by side (as fq_mul).
    static __device__ __forceinline__ void fq_mul_x3(const Fq& a1, const Fq& b1, Fq& r1, const Fq& a2,
                                                     const Fq& b2, Fq& r2, const Fq& a3, const Fq& b3,
                                                     Fq& r3) {
        uint32_t m1[9], m2[9], m3[9];
        uint64_t s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int j = 0; j < k; j++) {
                qmac(s1, a1.v[j], b1.v[k - j]);
                qmac(s2, a2.v[j], b2.v[k - j]);
                qmac(s3, a3.v[j], b3.v[k - j]);
                qmac(s1, m1[j], F::p(k - j));
                qmac(s2, m2[j], F::p(k - j));
                qmac(s3, m3[j], F::p(k - j));
            }
            qmac(s1, a1.v[k], b1.v[0]);
            qmac(s2, a2.v[k], b2.v[0]);
            qmac(s3, a3.v[k], b3.v[0]);
            m1[k] = ((uint32_t)s1 * F::kPinv) & ZQ_MASK;
            m2[k] = ((uint32_t)s2 * F::kPinv) & ZQ_MASK;
            m3[k] = ((uint32_t)s3 * F::kPinv) & ZQ_MASK;
            qmac(s1, m1[k], F::p(0));
            qmac(s2, m2[k], F::p(0));
            qmac(s3, m3[k], F::p(0));
            s1 >>= 29;
            s2 >>= 29;
            s3 >>= 29;
            carry_fence(s1);
            carry_fence(s2);
            carry_fence(s3);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
#pragma unroll
            for (int j = k - 8; j < 9; j++) {
                qmac(s1, a1.v[j], b1.v[k - j]);
                qmac(s2, a2.v[j], b2.v[k - j]);
                qmac(s3, a3.v[j], b3.v[k - j]);
                qmac(s1, m1[j], F::p(k - j));
                qmac(s2, m2[j], F::p(k - j));
                qmac(s3, m3[j], F::p(k - j));
            }
            r1.v[k - 9] = (uint32_t)s1 & ZQ_MASK;
            r2.v[k - 9] = (uint32_t)s2 & ZQ_MASK;
            r3.v[k - 9] = (uint32_t)s3 & ZQ_MASK;
            s1 >>= 29;
            s2 >>= 29;
            s3 >>= 29;
            carry_fence(s1);
            carry_fence(s2);
            carry_fence(s3);
        }
        r1.v[8] = (uint32_t)s1;
        r2.v[8] = (uint32_t)s2;
        r3.v[8] = (uint32_t)s3;
    }

    // r1 = a1^2, r2 = a2^2 side by side (as fq_sqr)
    static __device__ __forceinline__ void fq_sqr_x2(const Fq& a1, Fq& r1, const Fq& a2, Fq& r2) {
        uint32_t d1[9], d2[9], m1[9], m2[9];
#pragma unroll
        for (int i = 0; i < 9; i++) {
            d1[i] = a1.v[i] << 1;
            d2[i] = a2.v[i] << 1;
        }
        uint64_t s1 = 0, s2 = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int i = 0; 2 * i < k; i++) {
                qmac(s1, d1[i], a1.v[k - i]);
                qmac(s2, d2[i], a2.v[k - i]);
            }
            if ((k & 1) == 0) {
                qmac(s1, a1.v[k / 2], a1.v[k / 2]);
                qmac(s2, a2.v[k / 2], a2.v[k / 2]);
            }
#pragma unroll
            for (int j = 0; j < k; j++) {
                qmac(s1, m1[j], F::p(k - j));
                qmac(s2, m2[j], F::p(k - j));
            }
            m1[k] = ((uint32_t)s1 * F::kPinv) & ZQ_MASK;
            m2[k] = ((uint32_t)s2 * F::kPinv) & ZQ_MASK;
            qmac(s1, m1[k], F::p(0));
            qmac(s2, m2[k], F::p(0));
            s1 >>= 29;
            s2 >>= 29;
            carry_fence(s1);
            carry_fence(s2);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
#pragma unroll
            for (int i = k - 8; 2 * i < k; i++) {
                qmac(s1, d1[i], a1.v[k - i]);
                qmac(s2, d2[i], a2.v[k - i]);
            }
            if ((k & 1) == 0) {
                qmac(s1, a1.v[k / 2], a1.v[k / 2]);
                qmac(s2, a2.v[k / 2], a2.v[k / 2]);
            }
#pragma unroll
            for (int j = k - 8; j < 9; j++) {
                qmac(s1, m1[j], F::p(k - j));
                qmac(s2, m2[j], F::p(k - j));
            }
            r1.v[k - 9] = (uint32_t)s1 & ZQ_MASK;
            r2.v[k - 9] = (uint32_t)s2 & ZQ_MASK;
            s1 >>= 29;
            s2 >>= 29;
            carry_fence(s1);
            carry_fence(s2);
        }
        r1.v[8] = (uint32_t)s1;
        r2.v[8] = (uint32_t)s2;
    }

    // Two merged products side by side (four multiply-add chains):
    // r1 = fq_mul_add_sub<K1, B1>(a1, b1, c1, d1, e1), r2 likewise.
    template <int K1, int B1, int K2, int B2>
    static __device__ __forceinline__ void fq_mul_add_sub_x2(const Fq& a1, const Fq& b1, const Fq& c1,
                                                             const Fq& d1, const Fq& e1, Fq& r1,
                                                             const Fq& a2, const Fq& b2, const Fq& c2,
                                                             const Fq& d2, const Fq& e2, Fq& r2) {
        uint32_t m1[9], m2[9];
        uint64_t s1 = 0, t1 = 0, s2 = 0, t2 = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int j = 0; j < k; j++) {
                qmac(s1, a1.v[j], b1.v[k - j]);
                qmac(s2, a2.v[j], b2.v[k - j]);
                qmac(t1, c1.v[j], d1.v[k - j]);
                qmac(t2, c2.v[j], d2.v[k - j]);
                qmac((j & 1) ? t1 : s1, m1[j], F::p(k - j));
                qmac((j & 1) ? t2 : s2, m2[j], F::p(k - j));
            }
            qmac(s1, a1.v[k], b1.v[0]);
            qmac(s2, a2.v[k], b2.v[0]);
            qmac(t1, c1.v[k], d1.v[0]);
            qmac(t2, c2.v[k], d2.v[0]);
            s1 += t1;
            s2 += t2;
            m1[k] = ((uint32_t)s1 * F::kPinv) & ZQ_MASK;
            m2[k] = ((uint32_t)s2 * F::kPinv) & ZQ_MASK;
            qmac(s1, m1[k], F::p(0));
            qmac(s2, m2[k], F::p(0));
            s1 >>= 29;
            s2 >>= 29;
            t1 = 0;
            t2 = 0;
            carry_fence(s1);
            carry_fence(s2);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
            if (K1 > 0) s1 += q_kpr<K1, B1>(k - 9) - e1.v[k - 9];
            if (K2 > 0) s2 += q_kpr<K2, B2>(k - 9) - e2.v[k - 9];
#pragma unroll
            for (int j = k - 8; j < 9; j++) {
                qmac(s1, a1.v[j], b1.v[k - j]);
                qmac(s2, a2.v[j], b2.v[k - j]);
                qmac(t1, c1.v[j], d1.v[k - j]);
                qmac(t2, c2.v[j], d2.v[k - j]);
                qmac((j & 1) ? t1 : s1, m1[j], F::p(k - j));
                qmac((j & 1) ? t2 : s2, m2[j], F::p(k - j));
            }
            s1 += t1;
            s2 += t2;
            r1.v[k - 9] = (uint32_t)s1 & ZQ_MASK;
            r2.v[k - 9] = (uint32_t)s2 & ZQ_MASK;
            s1 >>= 29;
            s2 >>= 29;
            t1 = 0;
            t2 = 0;
            carry_fence(s1);
            carry_fence(s2);
        }
        r1.v[8] = (uint32_t)s1 + (K1 > 0 ? q_kpr<K1, B1>(8) - e1.v[8] : 0u);
        r2.v[8] = (uint32_t)s2 + (K2 > 0 ? q_kpr<K2, B2>(8) - e2.v[8] : 0u);
    }

    // (a b + c d) / 2^261 mod p with one Montgomery reduction: result <
    // (ab + cd)/2^261 + p, normalised.  a, c normalised; b limbs < 1.5 * 2^30,
    // d limbs < 2^30 (column sums < 2^63.75); ab + cd < p * 2^261 / 2.  Each
    // column's products go to two accumulators (a b and half the m p in one,
    // c d and the rest in the other), joined before m is taken.  With K > 0
    // (fq_mul_add_sub_x2) the limbs of K p - e enter the high columns as in
    // fq_mul_sub: result < (ab + cd)/2^261 + p + K p - e.
    static __device__ __forceinline__ Fq fq_mul_add(const Fq& a, const Fq& b, const Fq& c, const Fq& d) {
        uint32_t m[9];
        Fq t;
        uint64_t s1 = 0, s2 = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
#pragma unroll
            for (int j = 0; j < k; j++) {
                qmac(s1, a.v[j], b.v[k - j]);
                qmac(s2, c.v[j], d.v[k - j]);
                qmac((j & 1) ? s2 : s1, m[j], F::p(k - j));
            }
            qmac(s1, a.v[k], b.v[0]);
            qmac(s2, c.v[k], d.v[0]);
            s1 += s2;
            m[k] = ((uint32_t)s1 * F::kPinv) & ZQ_MASK;
            qmac(s1, m[k], F::p(0));
            s1 >>= 29;
            s2 = 0;
            carry_fence(s1);
        }
#pragma unroll
        for (int k = 9; k < 17; k++) {
#pragma unroll
            for (int j = k - 8; j < 9; j++) {
                qmac(s1, a.v[j], b.v[k - j]);
                qmac(s2, c.v[j], d.v[k - j]);
                qmac((j & 1) ? s2 : s1, m[j], F::p(k - j));
            }
            s1 += s2;
            t.v[k - 9] = (uint32_t)s1 & ZQ_MASK;
            s1 >>= 29;
            s2 = 0;
            carry_fence(s1);
        }
        t.v[8] = (uint32_t)s1;
        return t;
    }
    // x * 2^256 mod p (the 32-bit-limb Montgomery form) -> x * 2^261 mod p
    static __device__ __forceinline__ Fq fq_from_fp(const Fp& a) { return fq_mul(fq_slice(a), F::from32()); }

    // value < 8p with normalised limbs -> [0, p) in the 32-bit form (product < 1.1p)
    static __device__ __forceinline__ Fp fq_to_fp(const Fq& a) {
        return reduce_once32<F>(fq_pack(fq_mul(a, F::to32())));
    }

    // R mod p, R = 2^261
    static __device__ __forceinline__ Fq fq_one() { return F::one(); }

    // a^(p-2) = a^-1 for a != 0 mod p (0 -> 0); a < 8p with normalised limbs.
    // Left-to-right binary exponentiation over the 32-bit words of p - 2:
    // 256 squarings and one multiply per set bit.  Result < 1.5p.
    static __device__ Fq fq_inv(const Fq& a) {
        Fq r = fq_one();
#pragma unroll 1
        for (int w = 7; w >= 0; w--) {
            const uint32_t e = F::p32(w) - (w == 0 ? 2u : 0u);  // p32(0) >= 2
#pragma unroll 1
            for (int b = 31; b >= 0; b--) {
                r = fq_sqr(r);
                if ((e >> b) & 1u) r = fq_mul(r, a);
            }
        }
        return r;
    }

    // beta * R mod p (phi(x, y) = (beta x, y)); BN254 G1 only
    static __device__ __forceinline__ Fq fq_beta() { return F::beta(); }

    // a + b, normalised limbs (no reduction: the bound is the sum of the bounds)
    static __device__ __forceinline__ Fq fq_add(const Fq& a, const Fq& b) {
        Fq r;
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const uint32_t s = a.v[i] + b.v[i] + c;
            r.v[i] = s & ZQ_MASK;
            c = s >> 29;
        }
        return r;
    }

    // a - b + K*p for b < K*p: in [a - b + (K p - b), a + K p)
    template <int K>
    static __device__ __forceinline__ Fq fq_sub(const Fq& a, const Fq& b) {
        Fq r;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int32_t s = (int32_t)(a.v[i] + q_kp<K>(i)) - (int32_t)b.v[i] + c;
            r.v[i] = (uint32_t)s & ZQ_MASK;
            c = s >> 29;  // arithmetic: -1, 0 or 1
        }
        return r;
    }

    // a >= K*p ? a - K*p : a
    template <int K>
    static __device__ __forceinline__ Fq fq_csub(const Fq& a) {
        Fq d;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int32_t s = (int32_t)a.v[i] - (int32_t)q_kp<K>(i) + c;
            d.v[i] = (uint32_t)s & ZQ_MASK;
            c = s >> 29;
        }
        const bool keep = c < 0;
        Fq r;
#pragma unroll
        for (int i = 0; i < 9; i++) r.v[i] = keep ? a.v[i] : d.v[i];
        return r;
    }

    // a == 0 mod p for a < 2p.  The empty asm keeps the two OR-reductions whole:
    // without it the compiler splits "z == 0" into one compare per limb.
    static __device__ __forceinline__ bool fq_is_zero2(const Fq& a) {
        uint32_t z = 0, e = 0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            z |= a.v[i];
            e |= a.v[i] ^ F::p(i);
        }
        asm("" : "+v"(z), "+v"(e));
        return z == 0 || e == 0;
    }

    static __device__ __forceinline__ bool qxyzz_is_id(const QXyzz& p) { return fq_is_zero2(p.ZZ); }

    // doubling core: X, Y < 4p.  Returns X3 < 2.5p, Y3 < 3.5p; V, W < 1.5p.
    // (independent products in pairs: the doubling also runs alone in one
    // lane, in the device tail, where a second chain hides the latency)
    static __device__ __forceinline__ void qdbl_core(const Fq& X, const Fq& Y, Fq& X3, Fq& Y3, Fq& V, Fq& W) {
        const Fq U = fq_add(Y, Y);  // < 8p
        Fq xx, S;
        fq_sqr_x2(U, V, X, xx);                                       // U^2, X^2
        fq_mul_sub_x2<0, 1, 0, 1>(U, V, U, W, X, V, X, S);            // U V, X V
        const Fq M = fq_add(fq_add(xx, xx), xx);                    // < 4.5p
        X3 = fq_csub<2>(fq_sub<3>(fq_sqr(M), fq_add(S, S)));        // < 4.5p -> < 2.5p
        Fq MS, WY;
        fq_mul_sub_x2<0, 1, 0, 1>(M, fq_sub<3>(S, X3), M, MS, W, Y, W, WY);
        Y3 = fq_sub<2>(MS, WY);                                     // < 3.5p
    }

    // 2*(x, y) for an affine, non-identity point
    static __device__ __forceinline__ QXyzz qxyzz_dbl_aff(const Fq& x, const Fq& y) {
        QXyzz r;
        qdbl_core(x, y, r.X, r.Y, r.ZZ, r.ZZZ);
        return r;
    }

    // 2*P; identity stays identity (ZZ3 = V*ZZ1 = 0 mod p)
    static __device__ __forceinline__ QXyzz qxyzz_dbl(const QXyzz& p) {
        QXyzz r;
        Fq V, W;
        qdbl_core(p.X, p.Y, r.X, r.Y, V, W);
        r.ZZ = fq_mul(V, p.ZZ);
        r.ZZZ = fq_mul(W, p.ZZZ);
        return r;
    }

    // ---- repeated doubling (the device tail's window combine).  A chain of
    // doublings runs in Jacobian coordinates (x = X/Z^2, y = Y/Z^3;
    // dbl-2009-l for a = 0: 2M + 5S against XYZZ's 6M + 3S), with the
    // independent products in pairs (one lane runs the chain alone, so a
    // second multiply-add chain fills the dependent-issue latency).
    // Bounds: X, Y, Z < 2p in and out; Z = 0 mod p is the identity.
    static __device__ __forceinline__ void qjac_dbl(Fq& X, Fq& Y, Fq& Z) {
        Fq A, B, C, T;
        fq_sqr_x2(X, A, Y, B);                                        // < 1.5p each
        const Fq XB = fq_add(X, B);                                   // < 3.5p
        fq_sqr_x2(B, C, XB, T);                                       // < 1.5p each
        const Fq D0 = fq_sub<4>(T, fq_add(A, C));                     // (X+B)^2 - A - C, < 5.5p
        Fq D = fq_add(D0, D0);                                        // < 11p
        D = fq_csub<2>(fq_csub<4>(fq_csub<8>(D)));                    // < 2p
        const Fq E = fq_add(fq_add(A, A), A);                         // 3 X^2, < 4.5p
        const Fq Y2 = fq_add(Y, Y);                                   // < 4p
        Fq EE, Z3;
        fq_mul_sub_x2<0, 1, 0, 1>(E, E, E, EE, Y2, Z, Y2, Z3);         // E^2, 2 Y Z: < 1.5p each
        Fq X3 = fq_csub<2>(fq_csub<4>(fq_sub<4>(EE, fq_add(D, D))));  // E^2 - 2D: < 5.5p -> < 2p
        Fq C8 = fq_add(C, C);
        C8 = fq_add(C8, C8);
        C8 = fq_add(C8, C8);                                          // 8C < 12p
        // E (D - X3) - 8C + 13p: < 14.5p -> < 2p
        Fq Y3 = fq_mul_sub<13, 1>(E, fq_sub<2>(D, X3), C8);
        Y3 = fq_csub<2>(fq_csub<4>(fq_csub<8>(Y3)));
        X = X3;
        Y = Y3;
        Z = Z3;
    }

    // 2^k P for an XYZZ point (X < 6.5p, Y < 4p, ZZ, ZZZ < 2p): to Jacobian
    // (X ZZ^2, Y ZZZ^2, ZZZ), k doublings, back (ZZ = Z^2, ZZZ = Z^3).  The
    // identity (ZZ = ZZZ = 0) stays the identity.  Out: X, Y < 2p, ZZ, ZZZ < 1.5p.
    static __device__ __forceinline__ QXyzz qxyzz_dbl_k(const QXyzz& p, uint32_t k) {
        Fq ZZ2, ZZZ2, X, Y;
        fq_sqr_x2(p.ZZ, ZZ2, p.ZZZ, ZZZ2);
        fq_mul_sub_x2<0, 1, 0, 1>(p.X, ZZ2, p.X, X, p.Y, ZZZ2, p.Y, Y);  // < 1.5p each
        Fq Z = p.ZZZ;                                                     // < 2p
#pragma unroll 1
        for (uint32_t i = 0; i < k; i++) qjac_dbl(X, Y, Z);
        QXyzz r;
        r.X = X;
        r.Y = Y;
        r.ZZ = fq_sqr(Z);
        r.ZZZ = fq_mul(r.ZZ, Z);
        return r;
    }

    // acc += (x2, y2); x2, y2 < p.  q_id: the affine point is (0,0).
    static __device__ __forceinline__ void qxyzz_madd(QXyzz& acc, const Fq& x2, const Fq& y2, bool q_id) {
        const bool p_id = qxyzz_is_id(acc);
        const Fq U2 = fq_mul(x2, acc.ZZ);
        const Fq S2 = fq_mul(y2, acc.ZZZ);
        const Fq P = fq_sub<4>(U2, acc.X);  // < 5.5p
        const Fq R = fq_sub<4>(S2, acc.Y);
        const Fq PP = fq_sqr(P);
        const Fq PPP = fq_mul(P, PP);
        const Fq Q = fq_mul(acc.X, PP);
        const Fq RR = fq_sqr(R);
        const bool dbl = fq_is_zero2(PP) && fq_is_zero2(RR) && !p_id && !q_id;
        QXyzz r;
        r.X = fq_csub<3>(fq_sub<5>(RR, fq_add(fq_add(Q, Q), PPP)));         // < 6.5p -> < 3.5p
        r.Y = fq_sub<2>(fq_mul(R, fq_sub<4>(Q, r.X)), fq_mul(acc.Y, PPP));  // < 3.5p
        r.ZZ = fq_mul(acc.ZZ, PP);
        r.ZZZ = fq_mul(acc.ZZZ, PPP);
        if (__builtin_expect(dbl, 0)) r = qxyzz_dbl_aff(x2, y2);
        // P == -Q falls out of the formulas: P = 0 gives ZZ3 = 0 mod p, the identity.
        QXyzz q;
        q.X = x2;
        q.Y = y2;
        q.ZZ = fq_one();
        q.ZZZ = q.ZZ;
        r = qxyzz_select(p_id, q, r);
        acc = qxyzz_select(q_id, acc, r);
    }

    // Placeholder for an accumulator that starts a bucket run: the value is never
    // used (the run's first addition selects the point itself), but ZZ = ZZZ = 1
    // and X = 0 keep P = x2 - X away from 0, so the first addition does not look
    // exceptional.
    static __device__ __forceinline__ QXyzz qxyzz_run_start() {
        QXyzz r;
        r.X = fq_zero();
        r.Y = fq_zero();
        r.ZZ = fq_one();
        r.ZZZ = r.ZZ;
        return r;
    }

    // acc += (x2, y2) in the bucket-accumulation loop.  acc_id says acc is the
    // identity (a run start, or a cancellation earlier in the run); (x2, y2) is
    // never the identity there.  X3 is left below 6.5p (no final subtraction):
    // every later use of X takes it as a multiplicand or as the subtrahend of a
    // K = 7 form.  Doubling (acc == q) and cancellation (acc == -q) both have
    // PP = 0 mod p; PP < 1.6p is then 0 or p, so its low limb is 0 or p_0.  That one-limb test screens every step; the exact tests run only in
    // lanes that pass it, and a wave branches to them only if one of its lanes
    // does.  After a cancellation acc holds ZZ = 0 mod p, the stored form of the
    // identity.
    static __device__ __forceinline__ void qxyzz_madd_run(QXyzz& acc, bool& acc_id, const Fq& x2,
                                                   const Fq& y2) {
        // the ten products in pairs of independent ones, and Y3's two under one
        // reduction: Y3 = (R (Q - X3 + 7p) + Y1 (2p - PPP)) / 2^261
        Fq P, R, PP, RR, PPP, Q;
        fq_mul_sub_x2<7, 1, 5, 1>(x2, acc.ZZ, acc.X, P, y2, acc.ZZZ, acc.Y, R);  // < 8.5p, < 6.5p
        fq_sqr_x2(P, PP, R, RR);                                                   // < 1.6p each
        fq_mul_sub_x2<0, 1, 0, 1>(P, PP, P, PPP, acc.X, PP, P, Q);                 // < 1.5p each
        QXyzz r;
        Fq T;  // 2Q + PPP < 4.5p, limbs sums of three normalised limbs
#pragma unroll
        for (int i = 0; i < 9; i++) T.v[i] = Q.v[i] + Q.v[i] + PPP.v[i];
        r.X = fq_sub<5>(RR, T);                                                     // < 6.3p
        r.Y = fq_mul_add(R, fq_sub_lazy<7>(Q, r.X), acc.Y, fq_sub_lazy<2>(fq_zero(), PPP));  // < 1.4p
        fq_mul_sub_x2<0, 1, 0, 1>(acc.ZZ, PP, P, r.ZZ, acc.ZZZ, PPP, P, r.ZZZ);
        bool cancel = false;
        if (__builtin_expect(PP.v[0] == 0u || PP.v[0] == F::p(0), 0)) {
            const bool z = !acc_id && fq_is_zero2(PP);
            const bool d = z && fq_is_zero2(RR);
            cancel = z && !d;
            if (d) r = qxyzz_dbl_aff(x2, y2);
        }
        QXyzz q;
        q.X = x2;
        q.Y = y2;
        q.ZZ = fq_one();
        q.ZZZ = q.ZZ;
        acc = qxyzz_select(acc_id, q, r);
        acc_id = cancel;
    }

    // acc += q (both XYZZ; X < 6.5p as stored by the accumulation loop,
    // Y < 4p, ZZ, ZZZ < 2p).  Products in independent pairs, U2 - U1 and
    // S2 - S1 folded into their products, Y3's two products under one
    // reduction (as in qxyzz_madd_run): X3 < 4p, Y3 < 1.5p, ZZ3, ZZZ3 < 1.5p.
    static __device__ __forceinline__ void qxyzz_add(QXyzz& acc, const QXyzz& q) {
        const bool p_id = qxyzz_is_id(acc);
        const bool q_id = qxyzz_is_id(q);
        Fq U1, S1, P, R, PP, RR, PPP, Q, ZZ12, ZZZ12;
        fq_mul_sub_x2<0, 1, 0, 1>(acc.X, q.ZZ, acc.X, U1, acc.Y, q.ZZZ, acc.X, S1);
        fq_mul_sub_x2<2, 1, 2, 1>(q.X, acc.ZZ, U1, P, q.Y, acc.ZZZ, S1, R);  // < 3.5p
        fq_sqr_x2(P, PP, R, RR);
        fq_mul_sub_x2<0, 1, 0, 1>(P, PP, P, PPP, U1, PP, P, Q);
        const bool dbl = fq_is_zero2(PP) && fq_is_zero2(RR) && !p_id && !q_id;
        QXyzz r;
        Fq T;  // 2Q + PPP, limbs sums of three normalised limbs
#pragma unroll
        for (int i = 0; i < 9; i++) T.v[i] = Q.v[i] + Q.v[i] + PPP.v[i];
        r.X = fq_csub<3>(fq_sub<5>(RR, T));
        r.Y = fq_mul_add(R, fq_sub_lazy<4>(Q, r.X), S1, fq_sub_lazy<2>(fq_zero(), PPP));
        fq_mul_sub_x2<0, 1, 0, 1>(acc.ZZ, q.ZZ, P, ZZ12, acc.ZZZ, q.ZZZ, P, ZZZ12);
        fq_mul_sub_x2<0, 1, 0, 1>(ZZ12, PP, P, r.ZZ, ZZZ12, PPP, P, r.ZZZ);
        if (__builtin_expect(dbl, 0)) r = qxyzz_dbl(acc);
        r = qxyzz_select(p_id, q, r);
        acc = qxyzz_select(q_id, acc, r);
    }

    // ---- conversions between the two forms (canonical 32-bit form: [0, p),
    // R = 2^256; internal: R = 2^261)
    static __device__ __forceinline__ Fq fq_from_fp_canon(const Fp& a) { return fq_csub<1>(fq_from_fp(a)); }

    static __device__ __forceinline__ Xyzz qxyzz_to_xyzz(const QXyzz& a) {
        Xyzz r;
        r.X = fq_to_fp(a.X);
        r.Y = fq_to_fp(a.Y);
        r.ZZ = fq_to_fp(a.ZZ);
        r.ZZZ = fq_to_fp(a.ZZZ);
        return r;
    }

    static __device__ __forceinline__ QXyzz qxyzz_from_xyzz(const Xyzz& a) {
        QXyzz r;
        r.X = fq_from_fp_canon(canon32<F>(a.X));
        r.Y = fq_from_fp_canon(canon32<F>(a.Y));
        r.ZZ = fq_from_fp_canon(canon32<F>(a.ZZ));
        r.ZZZ = fq_from_fp_canon(canon32<F>(a.ZZZ));
        return r;
    }

    static __device__ __forceinline__ void store_qtab(uint4* base, uint64_t idx, const QAff& a) {
        uint32_t w[32];
        const Fq ny = fq_csub<1>(fq_sub<1>(fq_zero(), a.y));  // p - y; 0 stays 0
#pragma unroll
        for (int i = 0; i < 8; i++) {
            w[i] = a.x.v[i];
            w[8 + i] = a.y.v[i];
            w[16 + i] = ny.v[i];
        }
        w[24] = a.x.v[8]; w[25] = a.y.v[8];
        w[26] = a.x.v[8]; w[27] = ny.v[8];
#pragma unroll
        for (int i = 28; i < 32; i++) w[i] = 0;
        uint4* p = base + idx * 8;
#pragma unroll
        for (int i = 0; i < 8; i++) p[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    }
};

}  // namespace zk
