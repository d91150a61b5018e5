// Host-side BN254 arithmetic for the serial tail of the MSM (window combine,
// affine normalisation) and for building small constant tables.
// 4 x 64-bit little-endian Montgomery limbs (same bytes as the device's
// 8 x 32-bit limbs), XYZZ points with identity ZZ == 0.
#pragma once
#include <stdint.h>
#include <string.h>

#include <vector>

namespace zkh {

typedef unsigned __int128 u128;

struct Fp {
    uint64_t v[4];
};

// The two moduli of the BN254 cycle: p (BN254 G1 coordinates) and r
// (Grumpkin coordinates); kInv = -m^-1 mod 2^64, one() = 2^256 mod m,
// r2() = 2^512 mod m.
struct ModP {
    static constexpr uint64_t kP[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL,
                                       0xb85045b68181585dULL, 0x30644e72e131a029ULL};
    static constexpr uint64_t kPInv = 0x87d20782e4866389ULL;
    static Fp one() {
        return Fp{{0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL,
                   0x0e0a77c19a07df2fULL}};
    }
    static Fp r2() {
        return Fp{{0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL,
                   0x06d89f71cab8351fULL}};
    }
};

struct ModR {
    static constexpr uint64_t kP[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                                       0xb85045b68181585dULL, 0x30644e72e131a029ULL};
    static constexpr uint64_t kPInv = 0xc2e1f593efffffffULL;
    static Fp one() {
        return Fp{{0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL,
                   0x0e0a77c19a07df2fULL}};
    }
    static Fp r2() {
        return Fp{{0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL,
                   0x0216d0b17f4e44a5ULL}};
    }
};

inline Fp zero() { return Fp{{0, 0, 0, 0}}; }
inline bool is_zero(const Fp& a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
inline bool eq(const Fp& a, const Fp& b) { return memcmp(a.v, b.v, 32) == 0; }

struct Xyzz {
    Fp X, Y, ZZ, ZZZ;
};

// Field and XYZZ arithmetic modulo Mod (host side, 4 x 64-bit Montgomery)
template <class Mod>
struct HostArith {
    static Fp one() { return Mod::one(); }
    static Fp r2() { return Mod::r2(); }
    static Fp sub_p_if_ge(const Fp& t, uint64_t hi) {
        Fp d;
        uint64_t br = 0;
        for (int i = 0; i < 4; i++) {
            u128 x = (u128)t.v[i] - Mod::kP[i] - br;
            d.v[i] = (uint64_t)x;
            br = (uint64_t)(x >> 127) & 1;
        }
        return (hi || !br) ? d : t;
    }

    static Fp mul(const Fp& a, const Fp& b) {
        uint64_t t[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 4; i++) {
            uint64_t c = 0;
            for (int j = 0; j < 4; j++) {
                u128 x = (u128)a.v[j] * b.v[i] + t[j] + c;
                t[j] = (uint64_t)x;
                c = (uint64_t)(x >> 64);
            }
            u128 x = (u128)t[4] + c;
            t[4] = (uint64_t)x;
            t[5] = (uint64_t)(x >> 64);
            uint64_t m = t[0] * Mod::kPInv;
            x = (u128)m * Mod::kP[0] + t[0];
            c = (uint64_t)(x >> 64);
            for (int j = 1; j < 4; j++) {
                x = (u128)m * Mod::kP[j] + t[j] + c;
                t[j - 1] = (uint64_t)x;
                c = (uint64_t)(x >> 64);
            }
            x = (u128)t[4] + c;
            t[3] = (uint64_t)x;
            t[4] = t[5] + (uint64_t)(x >> 64);
        }
        Fp r{{t[0], t[1], t[2], t[3]}};
        return sub_p_if_ge(r, t[4]);
    }

    static Fp add(const Fp& a, const Fp& b) {
        Fp s;
        uint64_t c = 0;
        for (int i = 0; i < 4; i++) {
            u128 x = (u128)a.v[i] + b.v[i] + c;
            s.v[i] = (uint64_t)x;
            c = (uint64_t)(x >> 64);
        }
        return sub_p_if_ge(s, c);
    }

    static Fp sub(const Fp& a, const Fp& b) {
        Fp d;
        uint64_t br = 0;
        for (int i = 0; i < 4; i++) {
            u128 x = (u128)a.v[i] - b.v[i] - br;
            d.v[i] = (uint64_t)x;
            br = (uint64_t)(x >> 127) & 1;
        }
        if (br) {
            uint64_t c = 0;
            for (int i = 0; i < 4; i++) {
                u128 x = (u128)d.v[i] + Mod::kP[i] + c;
                d.v[i] = (uint64_t)x;
                c = (uint64_t)(x >> 64);
            }
        }
        return d;
    }

    static Fp dbl(const Fp& a) { return add(a, a); }

    static Fp inv(const Fp& a) {
        uint64_t e[4] = {Mod::kP[0] - 2, Mod::kP[1], Mod::kP[2], Mod::kP[3]};
        Fp acc = Mod::one(), base = a;
        for (int w = 0; w < 4; w++)
            for (int b = 0; b < 64; b++) {
                if ((e[w] >> b) & 1) acc = mul(acc, base);
                base = mul(base, base);
            }
        return acc;
    }


    static Xyzz identity() { return Xyzz{zero(), zero(), zero(), zero()}; }
    static bool is_id(const Xyzz& p) { return is_zero(p.ZZ); }

    static Xyzz xdbl(const Xyzz& p) {
        if (is_id(p)) return p;
        Fp U = dbl(p.Y), V = mul(U, U), W = mul(U, V), S = mul(p.X, V);
        Fp xx = mul(p.X, p.X), M = add(dbl(xx), xx);
        Xyzz r;
        r.X = sub(mul(M, M), dbl(S));
        r.Y = sub(mul(M, sub(S, r.X)), mul(W, p.Y));
        r.ZZ = mul(V, p.ZZ);
        r.ZZZ = mul(W, p.ZZZ);
        return r;
    }

    static Xyzz xadd(const Xyzz& a, const Xyzz& b) {
        if (is_id(a)) return b;
        if (is_id(b)) return a;
        Fp U1 = mul(a.X, b.ZZ), U2 = mul(b.X, a.ZZ), S1 = mul(a.Y, b.ZZZ), S2 = mul(b.Y, a.ZZZ);
        Fp P = sub(U2, U1), R = sub(S2, S1);
        if (is_zero(P)) {
            if (is_zero(R)) return xdbl(a);
            return identity();
        }
        Fp PP = mul(P, P), PPP = mul(P, PP), Q = mul(U1, PP);
        Xyzz r;
        r.X = sub(sub(mul(R, R), PPP), dbl(Q));
        r.Y = sub(mul(R, sub(Q, r.X)), mul(S1, PPP));
        r.ZZ = mul(mul(a.ZZ, b.ZZ), PP);
        r.ZZZ = mul(mul(a.ZZZ, b.ZZZ), PPP);
        return r;
    }

    static Xyzz from_affine(const Fp& x, const Fp& y) {
        if (is_zero(x) && is_zero(y)) return identity();
        return Xyzz{x, y, Mod::one(), Mod::one()};
    }

    // (x, y) with (0,0) for the identity
    static void to_affine(const Xyzz& p, Fp& x, Fp& y) {
        if (is_id(p)) {
            x = zero();
            y = zero();
            return;
        }
        Fp t = inv(mul(p.ZZ, p.ZZZ));
        x = mul(p.X, mul(t, p.ZZZ));
        y = mul(p.Y, mul(t, p.ZZ));
    }

    // Batch normalisation: every point needs 1/ZZ and 1/ZZZ; one inversion.
    static void batch_to_affine(const std::vector<Xyzz>& pts, std::vector<Fp>& xs,
                                std::vector<Fp>& ys) {
        size_t n = pts.size();
        xs.assign(n, zero());
        ys.assign(n, zero());
        std::vector<Fp> pre(2 * n + 1);
        pre[0] = Mod::one();
        for (size_t i = 0; i < n; i++) {
            Fp zz = is_id(pts[i]) ? Mod::one() : pts[i].ZZ;
            Fp zzz = is_id(pts[i]) ? Mod::one() : pts[i].ZZZ;
            pre[2 * i + 1] = mul(pre[2 * i], zz);
            pre[2 * i + 2] = mul(pre[2 * i + 1], zzz);
        }
        Fp acc = inv(pre[2 * n]);
        for (size_t i = n; i-- > 0;) {
            Fp zz = is_id(pts[i]) ? Mod::one() : pts[i].ZZ;
            Fp zzz = is_id(pts[i]) ? Mod::one() : pts[i].ZZZ;
            Fp izzz = mul(acc, pre[2 * i + 1]);
            acc = mul(acc, zzz);
            Fp izz = mul(acc, pre[2 * i]);
            acc = mul(acc, zz);
            if (!is_id(pts[i])) {
                xs[i] = mul(pts[i].X, izz);
                ys[i] = mul(pts[i].Y, izzz);
            }
        }
    }
};

using HP = HostArith<ModP>;  // BN254 G1
using HR = HostArith<ModR>;  // Grumpkin

// BN254 G2: XYZZ over Fp2 = Fp[u]/(u^2 + 1), components in the 2^256
// Montgomery form of p (Sources/zkMetal/Curve/BN254Pairing.swift:11-80)
struct Fp2 {
    Fp c0, c1;
};

struct Xyzz2 {
    Fp2 X, Y, ZZ, ZZZ;
};

struct HostG2 {
    using H = HP;
    static Fp2 zero2() { return Fp2{zero(), zero()}; }
    static Fp2 one2() { return Fp2{H::one(), zero()}; }
    static bool is_zero2(const Fp2& a) { return is_zero(a.c0) && is_zero(a.c1); }
    static Fp2 add2(const Fp2& a, const Fp2& b) { return Fp2{H::add(a.c0, b.c0), H::add(a.c1, b.c1)}; }
    static Fp2 sub2(const Fp2& a, const Fp2& b) { return Fp2{H::sub(a.c0, b.c0), H::sub(a.c1, b.c1)}; }
    static Fp2 dbl2(const Fp2& a) { return add2(a, a); }
    static Fp2 mul2(const Fp2& a, const Fp2& b) {
        const Fp t0 = H::mul(a.c0, b.c0), t1 = H::mul(a.c1, b.c1);
        const Fp t2 = H::mul(H::add(a.c0, a.c1), H::add(b.c0, b.c1));
        return Fp2{H::sub(t0, t1), H::sub(t2, H::add(t0, t1))};
    }
    // conj(a) / (a0^2 + a1^2)
    static Fp2 inv2(const Fp2& a) {
        const Fp ni = H::inv(H::add(H::mul(a.c0, a.c0), H::mul(a.c1, a.c1)));
        return Fp2{H::mul(a.c0, ni), H::sub(zero(), H::mul(a.c1, ni))};
    }

    static Xyzz2 identity() { return Xyzz2{zero2(), zero2(), zero2(), zero2()}; }
    static bool is_id(const Xyzz2& p) { return is_zero2(p.ZZ); }

    static Xyzz2 xdbl(const Xyzz2& p) {
        if (is_id(p)) return p;
        const Fp2 U = dbl2(p.Y), V = mul2(U, U), W = mul2(U, V), S = mul2(p.X, V);
        const Fp2 xx = mul2(p.X, p.X), M = add2(dbl2(xx), xx);
        Xyzz2 r;
        r.X = sub2(mul2(M, M), dbl2(S));
        r.Y = sub2(mul2(M, sub2(S, r.X)), mul2(W, p.Y));
        r.ZZ = mul2(V, p.ZZ);
        r.ZZZ = mul2(W, p.ZZZ);
        return r;
    }

    static Xyzz2 xadd(const Xyzz2& a, const Xyzz2& b) {
        if (is_id(a)) return b;
        if (is_id(b)) return a;
        const Fp2 U1 = mul2(a.X, b.ZZ), U2 = mul2(b.X, a.ZZ);
        const Fp2 S1 = mul2(a.Y, b.ZZZ), S2 = mul2(b.Y, a.ZZZ);
        const Fp2 P = sub2(U2, U1), R = sub2(S2, S1);
        if (is_zero2(P)) {
            if (is_zero2(R)) return xdbl(a);
            return identity();
        }
        const Fp2 PP = mul2(P, P), PPP = mul2(P, PP), Q = mul2(U1, PP);
        Xyzz2 r;
        r.X = sub2(sub2(mul2(R, R), PPP), dbl2(Q));
        r.Y = sub2(mul2(R, sub2(Q, r.X)), mul2(S1, PPP));
        r.ZZ = mul2(mul2(a.ZZ, b.ZZ), PP);
        r.ZZZ = mul2(mul2(a.ZZZ, b.ZZZ), PPP);
        return r;
    }

    // (x, y); the identity gives (0, 0)
    static void to_affine(const Xyzz2& p, Fp2& x, Fp2& y) {
        if (is_id(p)) {
            x = zero2();
            y = zero2();
            return;
        }
        const Fp2 t = inv2(mul2(p.ZZ, p.ZZZ));
        x = mul2(p.X, mul2(t, p.ZZZ));
        y = mul2(p.Y, mul2(t, p.ZZ));
    }
};
}  // namespace zkh
