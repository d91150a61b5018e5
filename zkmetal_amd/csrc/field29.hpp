// BN254 base field Fp on gfx950 in nine 29-bit limbs (radix 2^29), Montgomery
// form with R = 2^261.
//
// Why 29 bits: a column of the FIPS Montgomery product sums at most 9 limb
// products of a*b and 9 of m*p, each < 2^58, so the column stays far below
// 2^64 and every limb product is a single v_mad_u64_u32 into a plain 64-bit
// accumulator -- no carry-out instruction per product as with 32-bit limbs.
// 162 multiply-adds per product instead of 136 multiply-adds + 136 add-with-
// carry; with 29-bit limbs up to 2^30 (two normalised values added limb-wise)
// the column bound is 9*2^60 + 9*2^58 + 2^36 < 2^64.
//
// Values are "lazy": a product of inputs below 8p is below 1.5p (R > 2^7 p),
// so there is no final subtraction; each caller states the bound of what it
// holds.  Limbs are normalised (< 2^29) except where noted.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field.hpp"

namespace zk {

struct Fq {
    uint32_t v[9];
};

#define ZQ_MASK 0x1fffffffu
#define ZQ_PINV 0x04866389u  // -p^-1 mod 2^29

__device__ __forceinline__ uint32_t q_p(int i) {
    switch (i) {
        case 0: return 0x187cfd47u;
        case 1: return 0x010460b6u;
        case 2: return 0x1c72a34fu;
        case 3: return 0x02d522d0u;
        case 4: return 0x1585d978u;
        case 5: return 0x02db40c0u;
        case 6: return 0x00a6e141u;
        case 7: return 0x0e5c2634u;
        default: return 0x0030644eu;
    }
}

// column accumulation: acc += x * y (one v_mad_u64_u32)
__device__ __forceinline__ void qmac(uint64_t& acc, uint32_t x, uint32_t y) {
    acc += (uint64_t)x * y;
}

// An empty asm on the column carry (no instruction is emitted).  Without it
// the compiler's reassociation sums each column's products from zero, apart
// from the carry, and adds the carry with a separate 64-bit add plus register
// moves; with it the carry opens the next column's multiply-add chain.
// Measured in the bucket accumulation: 7% fewer VALU instructions per mixed
// addition, 3% less time.
__device__ __forceinline__ void carry_fence(uint64_t& acc) { asm("" : "+v"(acc)); }

// limb i of K*p, normalised (folded at compile time inside unrolled loops)
template <int K>
__device__ __forceinline__ uint32_t q_kp(int i) {
    uint64_t c = 0;
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = (uint64_t)K * q_p(j) + c;
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
__device__ __forceinline__ uint32_t q_kpr(int i) {
    const uint32_t c = q_kp<K>(i);
    return i == 0 ? c + (uint32_t)B * (1u << 29)
                  : (i < 8 ? c + (uint32_t)B * (1u << 29) - (uint32_t)B : c - (uint32_t)B);
}

// Q - c + K*p limb by limb with no carry propagation (limbs < 2^29 + B 2^29,
// value < Q + K*p): a multiplicand for fq_mul_sub / fq_mul whose other operand
// has normalised limbs.  Q normalised; c as for q_kpr<K, B>.
template <int K, int B = 1>
__device__ __forceinline__ Fq fq_sub_lazy(const Fq& q, const Fq& c) {
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
__device__ __forceinline__ Fq fq_mul_sub(const Fq& a, const Fq& b, const Fq& c) {
    uint32_t m[9];
    Fq t;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
#pragma unroll
        for (int j = 0; j < k; j++) {
            qmac(acc, a.v[j], b.v[k - j]);
            qmac(acc, m[j], q_p(k - j));
        }
        qmac(acc, a.v[k], b.v[0]);
        m[k] = ((uint32_t)acc * ZQ_PINV) & ZQ_MASK;
        qmac(acc, m[k], q_p(0));
        acc >>= 29;
        carry_fence(acc);
    }
#pragma unroll
    for (int k = 9; k < 17; k++) {
        if (K > 0) acc += q_kpr<K, B>(k - 9) - c.v[k - 9];
#pragma unroll
        for (int j = k - 8; j < 9; j++) {
            qmac(acc, a.v[j], b.v[k - j]);
            qmac(acc, m[j], q_p(k - j));
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
__device__ __forceinline__ Fq fq_mul(const Fq& a, const Fq& b) { return fq_mul_sub<0>(a, b, a); }

// a^2 / 2^261 mod p: the symmetric products a_i a_j (i != j) once, from the
// doubled limb 2 a_i (< 2^30), plus the squares: 45 + 81 multiply-adds.
// Column bound: 4 doubled products < 2^59, one square < 2^58, 9 m*p products
// < 2^58 and the carry: < 18 * 2^58 < 2^63.  a: normalised limbs, value < 8p.
// fq_sqr_sub<K, B>(a, c) = a^2 / 2^261 - c + K*p, as fq_mul_sub.
template <int K, int B = 1>
__device__ __forceinline__ Fq fq_sqr_sub(const Fq& a, const Fq& c) {
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
        for (int j = 0; j < k; j++) qmac(acc, m[j], q_p(k - j));
        m[k] = ((uint32_t)acc * ZQ_PINV) & ZQ_MASK;
        qmac(acc, m[k], q_p(0));
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
        for (int j = k - 8; j < 9; j++) qmac(acc, m[j], q_p(k - j));
        t.v[k - 9] = (uint32_t)acc & ZQ_MASK;
        acc >>= 29;
        carry_fence(acc);
    }
    t.v[8] = (uint32_t)acc + (K > 0 ? q_kpr<K, B>(8) - c.v[8] : 0u);
    return t;
}

__device__ __forceinline__ Fq fq_sqr(const Fq& a) { return fq_sqr_sub<0>(a, a); }

// 256-bit integer (8 x 32) <-> 9 x 29 bit slices of the same integer
__device__ __forceinline__ Fq fq_slice(const Fp& a) {
    Fq r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int bit = 29 * i, w = bit >> 5, s = bit & 31;
        uint32_t lo = a.v[w] >> s;
        if (s > 3 && w + 1 < 8) lo |= a.v[w + 1] << (32 - s);
        r.v[i] = lo & ZQ_MASK;
    }
    return r;
}

__device__ __forceinline__ Fp fq_pack(const Fq& a) {
    Fp r;
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const int bit = 32 * w, i = bit / 29, s = bit % 29;
        uint32_t x = a.v[i] >> s;
        if (i + 1 < 9) x |= a.v[i + 1] << (29 - s);
        if (s > 26 && i + 2 < 9) x |= a.v[i + 2] << (58 - s);
        r.v[w] = x;
    }
    return r;
}

// x * 2^256 mod p (the 32-bit-limb Montgomery form) -> x * 2^261 mod p
__device__ __forceinline__ Fq fq_from_fp(const Fp& a) {
    Fq k;  // 2^5 * 2^261 mod p
    k.v[0] = 0x13349ca1u; k.v[1] = 0x1a5d84a8u; k.v[2] = 0x0a3e5cacu; k.v[3] = 0x100249e0u;
    k.v[4] = 0x12b951e8u; k.v[5] = 0x0e92d304u; k.v[6] = 0x14cb95b3u; k.v[7] = 0x041b9d3du;
    k.v[8] = 0x00058003u;
    return fq_mul(fq_slice(a), k);
}

// value < 8p with normalised limbs -> [0, p) in the 32-bit form
__device__ __forceinline__ Fp fq_to_fp(const Fq& a) {
    Fq k;  // 2^256 mod p
    k.v[0] = 0x058f0d9du; k.v[1] = 0x1aea1c6eu; k.v[2] = 0x11c2cf74u; k.v[3] = 0x11d651ebu;
    k.v[4] = 0x1462c0a7u; k.v[5] = 0x11b7bc3cu; k.v[6] = 0x1cbd99bau; k.v[7] = 0x183340fbu;
    k.v[8] = 0x000e0a77u;
    return fp_reduce_once(fq_pack(fq_mul(a, k)));  // product < 1.5p
}

}  // namespace zk
