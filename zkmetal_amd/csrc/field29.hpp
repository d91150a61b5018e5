// BN254's two prime fields on gfx950 in nine 29-bit limbs (radix 2^29),
// Montgomery form with R = 2^261; the arithmetic is Arith<F> (curve29.hpp).
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

// The two fields of the BN254 cycle, as constants for the 29-bit arithmetic
// (curve29.hpp, Arith<F>): p(i) the nine 29-bit limbs of the modulus m,
// kPinv = -m^-1 mod 2^29, p32(i) its eight 32-bit limbs, one() = 2^261 mod m,
// from32() = 2^266 mod m (a Montgomery product with it maps the 32-bit
// Montgomery form x 2^256 to x 2^261), to32() = 2^256 mod m (maps back).
// Both moduli are below 2^254, so the lazy bounds hold for either
// (2^261 > 2^7 m).

// BN254 base field: coordinates of BN254 G1
struct Bn254Fp {
    __device__ static __forceinline__ uint32_t p(int i) {
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
    static constexpr uint32_t kPinv = 0x04866389u;
    __device__ static __forceinline__ uint32_t p32(int i) { return p_limb(i); }
    __device__ static __forceinline__ Fq one() {
        return Fq{{0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u, 0x1d4240ceu,
                   0x11d54c07u, 0x052ac7a8u, 0x000dc836u}};
    }
    __device__ static __forceinline__ Fq from32() {
        return Fq{{0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u, 0x0e92d304u,
                   0x14cb95b3u, 0x041b9d3du, 0x00058003u}};
    }
    __device__ static __forceinline__ Fq to32() {
        return Fq{{0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u, 0x11b7bc3cu,
                   0x1cbd99bau, 0x183340fbu, 0x000e0a77u}};
    }
    // beta * 2^261 mod p, beta a cube root of unity: phi(x, y) = (beta x, y)
    __device__ static __forceinline__ Fq beta() {
        return Fq{{0x0a337995u, 0x158d1d23u, 0x189c9b98u, 0x12fa4e45u, 0x185faadcu, 0x0176f16du,
                   0x0eed93bau, 0x14291140u, 0x000c0afeu}};
    }
};

// BN254 scalar field r: coordinates of Grumpkin (y^2 = x^3 - 17 over Fr)
struct Bn254Fr {
    __device__ static __forceinline__ uint32_t p(int i) {
        switch (i) {
            case 0: return 0x10000001u;
            case 1: return 0x1f0fac9fu;
            case 2: return 0x0e5c2450u;
            case 3: return 0x07d090f3u;
            case 4: return 0x1585d283u;
            case 5: return 0x02db40c0u;
            case 6: return 0x00a6e141u;
            case 7: return 0x0e5c2634u;
            default: return 0x0030644eu;
        }
    }
    static constexpr uint32_t kPinv = 0x0fffffffu;
    __device__ static __forceinline__ uint32_t p32(int i) {
        switch (i) {
            case 0: return 0xf0000001u;
            case 1: return 0x43e1f593u;
            case 2: return 0x79b97091u;
            case 3: return 0x2833e848u;
            case 4: return 0x8181585du;
            case 5: return 0xb85045b6u;
            case 6: return 0xe131a029u;
            default: return 0x30644e72u;
        }
    }
    __device__ static __forceinline__ Fq one() {
        return Fq{{0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu, 0x1d4240ceu,
                   0x11d54c07u, 0x052ac7a8u, 0x000dc836u}};
    }
    __device__ static __forceinline__ Fq from32() {
        return Fq{{0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau, 0x0e92d304u,
                   0x14cb95b3u, 0x041b9d3du, 0x00058003u}};
    }
    __device__ static __forceinline__ Fq to32() {
        return Fq{{0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu, 0x11b7bc3cu,
                   0x1cbd99bau, 0x183340fbu, 0x000e0a77u}};
    }
};

// t < 2m (32-bit limbs) -> t mod m
template <class F>
__device__ __forceinline__ Fp reduce_once32(const Fp& t) {
    Fp d;
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(t.v[i], F::p32(i), br, &br);
    return br ? t : d;
}

// the canonical representative of a 256-bit value (at most 6 subtractions
// of m: 2^256 < 6 m)
template <class F>
__device__ __forceinline__ Fp canon32(Fp a) {
#pragma unroll 1
    for (int k = 0; k < 6; k++) {
        Fp d;
        unsigned br = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(a.v[i], F::p32(i), br, &br);
        if (br) break;
        a = d;
    }
    return a;
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

}  // namespace zk
