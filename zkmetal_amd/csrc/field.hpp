// BN254 base field Fp on gfx950: 8 x 32-bit little-endian limbs, Montgomery
// form with R = 2^256, values kept fully reduced in [0, p).
//
// Multiplication is FIPS (finely integrated product scanning) Montgomery:
// column k of a*b and of m*p is summed into a 96-bit accumulator, one
// v_mad_u64_u32 (32x32+64 -> 64 with carry-out) plus one v_addc_co_u32 per
// limb product, m_k = low(acc) * (-p^-1) mod 2^32.  That is 64 + 64 limb
// products and 8 v_mul_lo_u32 per field multiply, the same count as the
// reference's 8x32-bit CIOS (Sources/Shaders/fields/bn254_fp.metal:134-174)
// with the carry handling folded into the accumulator.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zk {

struct Fp {
    uint32_t v[8];
};

// p = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
#define ZK_P0 0xd87cfd47u
#define ZK_P1 0x3c208c16u
#define ZK_P2 0x6871ca8du
#define ZK_P3 0x97816a91u
#define ZK_P4 0x8181585du
#define ZK_P5 0xb85045b6u
#define ZK_P6 0xe131a029u
#define ZK_P7 0x30644e72u
#define ZK_PINV32 0xe4866389u  // -p^-1 mod 2^32

__device__ __forceinline__ uint32_t p_limb(int i) {
    switch (i) {
        case 0: return ZK_P0;
        case 1: return ZK_P1;
        case 2: return ZK_P2;
        case 3: return ZK_P3;
        case 4: return ZK_P4;
        case 5: return ZK_P5;
        case 6: return ZK_P6;
        default: return ZK_P7;
    }
}

// R mod p (Montgomery one)
__device__ __forceinline__ Fp fp_one() {
    Fp r;
    r.v[0] = 0xc58f0d9du; r.v[1] = 0xd35d438du; r.v[2] = 0xf5c70b3du; r.v[3] = 0x0a78eb28u;
    r.v[4] = 0x7879462cu; r.v[5] = 0x666ea36fu; r.v[6] = 0x9a07df2fu; r.v[7] = 0x0e0a77c1u;
    return r;
}

// beta (cube root of unity, Montgomery form): phi(x, y) = (beta x, y)
__device__ __forceinline__ Fp fp_beta() {
    Fp r;
    r.v[0] = 0xd782e155u; r.v[1] = 0x71930c11u; r.v[2] = 0xffbe3323u; r.v[3] = 0xa6bb947cu;
    r.v[4] = 0xd4741444u; r.v[5] = 0xaa303344u; r.v[6] = 0x26594943u; r.v[7] = 0x2c3b3f0du;
    return r;
}

__device__ __forceinline__ Fp fp_zero() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
}

// acc(96 bit: lo64, hi32) += x * y
__device__ __forceinline__ void mac(uint64_t& lo, uint32_t& hi, uint32_t x, uint32_t y) {
    uint64_t c;
    asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
        "v_addc_co_u32_e64 %2, %1, 0, %2, %1"
        : "+v"(lo), "=&s"(c), "+v"(hi)
        : "v"(x), "v"(y));
}

__device__ __forceinline__ void shift32(uint64_t& lo, uint32_t& hi) {
    lo = (lo >> 32) | ((uint64_t)hi << 32);
    hi = 0;
}

// r = t - p if t >= p else t   (t < 2p)
__device__ __forceinline__ Fp fp_reduce_once(const Fp& t) {
    Fp d;
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(t.v[i], p_limb(i), br, &br);
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = br ? t.v[i] : d.v[i];
    return r;
}

__device__ __forceinline__ Fp fp_mul(const Fp& a, const Fp& b) {
    uint32_t m[8];
    Fp t;
    uint64_t lo = 0;
    uint32_t hi = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int j = 0; j < i; j++) {
            mac(lo, hi, a.v[j], b.v[i - j]);
            mac(lo, hi, m[j], p_limb(i - j));
        }
        mac(lo, hi, a.v[i], b.v[0]);
        m[i] = (uint32_t)lo * ZK_PINV32;
        mac(lo, hi, m[i], ZK_P0);
        shift32(lo, hi);
    }
#pragma unroll
    for (int i = 8; i < 15; i++) {
#pragma unroll
        for (int j = i - 7; j < 8; j++) {
            mac(lo, hi, a.v[j], b.v[i - j]);
            mac(lo, hi, m[j], p_limb(i - j));
        }
        t.v[i - 8] = (uint32_t)lo;
        shift32(lo, hi);
    }
    t.v[7] = (uint32_t)lo;
    return fp_reduce_once(t);
}

__device__ __forceinline__ Fp fp_sqr(const Fp& a) { return fp_mul(a, a); }

__device__ __forceinline__ Fp fp_add(const Fp& a, const Fp& b) {
    Fp t;
    unsigned c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = __builtin_addc(a.v[i], b.v[i], c, &c);
    return fp_reduce_once(t);  // a + b < 2p < 2^256: no carry out of limb 7
}

__device__ __forceinline__ Fp fp_sub(const Fp& a, const Fp& b) {
    Fp t;
    unsigned br = 0, c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = __builtin_subc(a.v[i], b.v[i], br, &br);
    uint32_t mask = 0u - br;
#pragma unroll
    for (int i = 0; i < 8; i++) t.v[i] = __builtin_addc(t.v[i], p_limb(i) & mask, c, &c);
    return t;
}

__device__ __forceinline__ Fp fp_dbl(const Fp& a) { return fp_add(a, a); }

__device__ __forceinline__ bool fp_is_zero(const Fp& a) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) x |= a.v[i];
    return x == 0;
}

// -a (0 stays 0)
__device__ __forceinline__ Fp fp_neg(const Fp& a) {
    Fp r;
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_subc(p_limb(i), a.v[i], br, &br);
    bool z = fp_is_zero(a);
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = z ? 0u : r.v[i];
    return r;
}

__device__ __forceinline__ Fp fp_select(bool c, const Fp& a, const Fp& b) {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = c ? a.v[i] : b.v[i];
    return r;
}

// Reduce an arbitrary 256-bit value into [0, p) (at most 5 subtractions).
__device__ __forceinline__ Fp fp_canon(Fp a) {
#pragma unroll 1
    for (int k = 0; k < 6; k++) {
        Fp d;
        unsigned br = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(a.v[i], p_limb(i), br, &br);
        if (br) break;
        a = d;
    }
    return a;
}

// a^(p-2) by a fixed square-and-multiply chain over the exponent bits.
__device__ inline Fp fp_inv(const Fp& a) {
    const uint32_t e[8] = {ZK_P0 - 2u, ZK_P1, ZK_P2, ZK_P3, ZK_P4, ZK_P5, ZK_P6, ZK_P7};
    Fp acc = fp_one();
    Fp base = a;
#pragma unroll 1
    for (int w = 0; w < 8; w++) {
        uint32_t bits = e[w];
#pragma unroll 1
        for (int b = 0; b < 32; b++) {
            if (bits & 1u) acc = fp_mul(acc, base);
            base = fp_sqr(base);
            bits >>= 1;
        }
    }
    return acc;
}

}  // namespace zk
