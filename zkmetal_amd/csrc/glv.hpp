// Scalar side of the MSM on the device: reduction mod r, the GLV split
// k == (+-k1) + lambda*(+-k2) (mod r) of Sources/zkMetal/MSM/GLV.swift:37-92 /
// Sources/Shaders/msm/glv_kernels.metal:172-256 (same lattice constants and
// truncated Babai rounding), and the signed c-bit recoding of
// Sources/Shaders/msm/msm_kernels.metal:399-439.
//
// Bounds (checked numerically from the constants): c1*a2 < r, so one
// conditional "+r" after each subtraction keeps k1 in [0, r); the folded
// |k1| and |k2| are below 2^126.97, so ceil(128/c) signed windows never leave
// a final carry (the kernel still reports one as an error).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zk {

// r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
__device__ __forceinline__ uint32_t r_limb(int i) {
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

// Fr element in Montgomery form (x * 2^256 mod r, the reference's Fr limbs)
// -> standard form x, for any 256-bit input: Montgomery reduction REDC(k),
// the same value as the reference's fr_mul(k, 1)
// (Sources/NeonFieldOps/bn254_msm.c:1537-1552).  The REDC result is below
// r + 1, so one conditional subtraction makes it canonical.
#define ZK_RINV32 0xefffffffu  // -r^-1 mod 2^32
__device__ __forceinline__ void fr_from_mont(uint32_t k[8]) {
    uint32_t t[17];
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = k[i];
#pragma unroll
    for (int i = 8; i < 17; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t m = t[i] * ZK_RINV32;  // t + m r == 0 mod 2^(32(i+1))
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint64_t s = (uint64_t)m * r_limb(j) + t[i + j] + c;
            t[i + j] = (uint32_t)s;
            c = s >> 32;
        }
#pragma unroll
        for (int j = i + 8; j < 17; j++) {
            const uint64_t s = (uint64_t)t[j] + c;
            t[j] = (uint32_t)s;
            c = s >> 32;
        }
    }
    uint32_t d[8];
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d[i] = __builtin_subc(t[8 + i], r_limb(i), br, &br);
    br = (t[16] == 0) ? br : 0u;  // t[16] is 0: the result is < r + 1 < 2^256
#pragma unroll
    for (int i = 0; i < 8; i++) k[i] = br ? t[8 + i] : d[i];
}

// out[0..na+nb) = a[0..na) * b[0..nb)
template <int NA, int NB>
__device__ __forceinline__ void mul_wide(const uint32_t* a, const uint32_t* b, uint32_t* out) {
#pragma unroll
    for (int i = 0; i < NA + NB; i++) out[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        uint32_t carry = 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            uint64_t t = (uint64_t)a[i] * b[j] + out[i + j] + carry;
            out[i + j] = (uint32_t)t;
            carry = (uint32_t)(t >> 32);
        }
        out[i + NB] = carry;
    }
}

// a -= b over N limbs; returns the borrow
template <int N>
__device__ __forceinline__ unsigned sub_n(uint32_t* a, const uint32_t* b) {
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < N; i++) a[i] = __builtin_subc(a[i], b[i], br, &br);
    return br;
}

__device__ __forceinline__ void add_r(uint32_t* a) {
    unsigned c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = __builtin_addc(a[i], r_limb(i), c, &c);
}

__device__ __forceinline__ bool geq_r(const uint32_t* a) {
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) (void)__builtin_subc(a[i], r_limb(i), br, &br);
    return br == 0;
}

struct GlvSplit {
    uint32_t k1[4], k2[4];
    bool neg1, neg2;
};

__device__ inline GlvSplit glv_split(const uint32_t kin[8]) {
    uint32_t k[8];
#pragma unroll
    for (int i = 0; i < 8; i++) k[i] = kin[i];
#pragma unroll 1
    for (int it = 0; it < 6 && geq_r(k); it++) {
        uint32_t rr[8];
#pragma unroll
        for (int i = 0; i < 8; i++) rr[i] = r_limb(i);
        sub_n<8>(k, rr);
    }
    const uint32_t g1[5] = {0x391eb18du, 0x7a7bd9d4u, 0xa773d2cfu, 0x4ccef014u, 0x2u};
    const uint32_t g2[3] = {0xc7e0b3d7u, 0xd91d232eu, 0x2u};
    const uint32_t a1[2] = {0x94d213e3u, 0x89d32568u};
    const uint32_t a2[4] = {0x1221250bu, 0x0be4e154u, 0xeeb859fdu, 0x6f4d8248u};
    const uint32_t mb1[4] = {0x7d4f1128u, 0x8211bbebu, 0xeeb859fcu, 0x6f4d8248u};
    uint32_t w13[13], w11[11];
    mul_wide<8, 5>(k, g1, w13);
    mul_wide<8, 3>(k, g2, w11);
    uint32_t c1[4] = {w13[8], w13[9], w13[10], w13[11]};  // floor(k g1 / 2^256)
    uint32_t c2[2] = {w11[8], w11[9]};                    // floor(k g2 / 2^256)

    // k1 = k - c2 a1 - c1 a2 (mod r)
    uint32_t t4[4], t8[8];
    mul_wide<2, 2>(c2, a1, t4);
    uint32_t s8[8] = {t4[0], t4[1], t4[2], t4[3], 0, 0, 0, 0};
    uint32_t k1[8];
#pragma unroll
    for (int i = 0; i < 8; i++) k1[i] = k[i];
    if (sub_n<8>(k1, s8)) add_r(k1);
    mul_wide<4, 4>(c1, a2, t8);
    if (sub_n<8>(k1, t8)) add_r(k1);

    // k2 = c2 (-b1) - c1 b2 as a signed integer
    uint32_t x6[6], y6[6];
    mul_wide<2, 4>(c2, mb1, x6);
    mul_wide<4, 2>(c1, a1, y6);  // b2 == a1
    bool n2 = sub_n<6>(x6, y6) != 0;
    if (n2) {
        unsigned br = 0;
#pragma unroll
        for (int i = 0; i < 6; i++) x6[i] = __builtin_subc(0u, x6[i], br, &br);
    }
    // fold k1 into [0, r/2]
    const uint32_t half[8] = {0xf8000000u, 0xa1f0fac9u, 0x3cdcb848u, 0x9419f424u,
                              0x40c0ac2eu, 0xdc2822dbu, 0x7098d014u, 0x18322739u};
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) (void)__builtin_subc(k1[i], half[i], br, &br);
    bool n1 = (br == 0);
    if (n1) {
        uint32_t rr[8];
#pragma unroll
        for (int i = 0; i < 8; i++) rr[i] = r_limb(i);
        sub_n<8>(rr, k1);
#pragma unroll
        for (int i = 0; i < 8; i++) k1[i] = rr[i];
    }
    GlvSplit s;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s.k1[i] = k1[i];
        s.k2[i] = x6[i];
    }
    s.neg1 = n1;
    s.neg2 = n2;
    return s;
}

// c-bit window w of a 128-bit value held in 4 little-endian limbs
// bits [w c, w c + c) of the 128-bit v, from two 64-bit halves by shifts (an
// indexed register array would be placed in scratch memory)
__device__ __forceinline__ uint32_t window_bits128(const uint32_t v[4], uint32_t w, uint32_t c) {
    const uint64_t lo = ((uint64_t)v[1] << 32) | v[0], hi = ((uint64_t)v[3] << 32) | v[2];
    const uint32_t off = w * c;
    uint64_t x;
    if (off >= 128u) x = 0;
    else if (off >= 64u) x = hi >> (off - 64u);
    else x = (lo >> off) | (off ? hi << (64u - off) : 0ull);
    return (uint32_t)x & ((1u << c) - 1u);
}

// Digit word layout (one per window per half-scalar):
//   bits 0..15  bucket key = |d| - 1, or 0xFFFF when d == 0 (no bucket)
//   bit  31     the point is subtracted (digit sign xor GLV sign)
#define ZK_KEY_NONE 0xFFFFu

}  // namespace zk
