// BN254 G2 XYZZ arithmetic: the twist y^2 = x^3 + 3/(9+u) over
// Fp2 = Fp[u]/(u^2 + 1) (Sources/zkMetal/Curve/BN254Pairing.swift:11-80,
// 300-372), on the 29-bit Fp arithmetic of the MSM kernels (Arith<Bn254Fp>).
// The formulas are those of the G1 kernels (madd-2008-s, add-2008-s,
// dbl-2008-s-1) with Fp2 coefficients; b' enters none of them.
//
// Bounds: every Fp2 value held between operations has both components below
// 2p with normalised limbs (so fq_is_zero2 decides == 0, and sums of two stay
// below 4p as multiplicands).  Each operation restores that with conditional
// subtractions: a product of components < 4p is < 1.1p; mul2's c0 = t0 - t1
// + 2p < 3.1p and c1 = t2 - (t0 + t1) + 3p < 4.1p, sub2 = a - b + 2p < 4p,
// 2a < 4p, 3a < 6p are brought under 2p by one subtraction of 2p (below 4p)
// or of 4p then 2p (below 8p).
// Stored form: a point is two 144-byte halves (the c0 limbs of X, Y, ZZ, ZZZ,
// then the c1 limbs) and a table point two 128-byte slots (x.c0, y.c0,
// (p - y).c0, then the c1 parts), so the G1 loads and stores serve both.
#pragma once
#include "curve29.hpp"

namespace zk {

struct Fq2 {
    Fq c0, c1;
};

struct QAff2 {
    Fq2 x, y;
};

struct QXyzz2 {
    Fq2 X, Y, ZZ, ZZZ;
};

struct G2Arith {
    using A = Arith<Bn254Fp>;

    // a < 4p -> a < 2p;  a < 8p -> a < 2p
    static __device__ __forceinline__ Fq red4(const Fq& a) { return A::fq_csub<2>(a); }
    static __device__ __forceinline__ Fq red8(const Fq& a) { return A::fq_csub<2>(A::fq_csub<4>(a)); }

    static __device__ __forceinline__ Fq2 zero2() { return Fq2{fq_zero(), fq_zero()}; }
    static __device__ __forceinline__ Fq2 one2() { return Fq2{A::fq_one(), fq_zero()}; }

    static __device__ __forceinline__ Fq2 sel2(bool c, const Fq2& a, const Fq2& b) {
        return Fq2{fq_select(c, a.c0, b.c0), fq_select(c, a.c1, b.c1)};
    }

    static __device__ __forceinline__ bool is_zero2(const Fq2& a) {
        return A::fq_is_zero2(a.c0) & A::fq_is_zero2(a.c1);
    }

    // Karatsuba: (a0 b0 - a1 b1) + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) u; the
    // three products side by side (independent multiply-add chains: one wave
    // per SIMD runs this code, so its own chains must hide each other's
    // latency)
    static __device__ __forceinline__ Fq2 mul2(const Fq2& a, const Fq2& b) {
        Fq t0, t1, t2;
        A::fq_mul_x3(a.c0, b.c0, t0, a.c1, b.c1, t1, A::fq_add(a.c0, a.c1), A::fq_add(b.c0, b.c1), t2);
        return Fq2{red4(A::fq_sub<2>(t0, t1)), red8(A::fq_sub<3>(t2, A::fq_add(t0, t1)))};
    }

    // (a0 + a1)(a0 - a1) + 2 a0 a1 u, the two products side by side
    static __device__ __forceinline__ Fq2 sqr2(const Fq2& a) {
        Fq c0, t;
        A::fq_mul_sub_x2<0, 1, 0, 1>(A::fq_add(a.c0, a.c1), A::fq_sub<2>(a.c0, a.c1), a.c0, c0, a.c0, a.c1,
                                     a.c0, t);
        return Fq2{c0, red4(A::fq_add(t, t))};
    }

    static __device__ __forceinline__ Fq2 sub2(const Fq2& a, const Fq2& b) {
        return Fq2{red4(A::fq_sub<2>(a.c0, b.c0)), red4(A::fq_sub<2>(a.c1, b.c1))};
    }

    static __device__ __forceinline__ Fq2 dbl2(const Fq2& a) {
        return Fq2{red4(A::fq_add(a.c0, a.c0)), red4(A::fq_add(a.c1, a.c1))};
    }

    // ---- the bucket loop's products: schoolbook Fp2 with merged reductions,
    // c0 = (a0 b0 + a1 (KB p - b1)) / 2^261, c1 = (a0 b1 + a1 b0) / 2^261
    // (the same 4 products of 81 multiply-adds and 2 reductions as Karatsuba's
    // 3 products, with normalised outputs and no conditional subtraction),
    // each component minus e + K p folded into its high columns.  a, b
    // components normalised, b1 < KB p; results < (|a| |b| + |a| KB p)/2^261
    // + p + K p - e.
    template <int K>
    static __device__ __forceinline__ Fq nl(const Fq& v) { return A::fq_sub_lazy<K>(fq_zero(), v); }

    template <int K, int KB = 2>
    static __device__ __forceinline__ Fq2 mul2m(const Fq2& a, const Fq2& b, const Fq2& e) {
        Fq2 r;
        A::fq_mul_add_sub_x2<K, 1, K, 1>(a.c0, b.c0, a.c1, nl<KB>(b.c1), e.c0, r.c0, a.c0, b.c1, a.c1, b.c0,
                                         e.c1, r.c1);
        return r;
    }

    static __device__ __forceinline__ Fq lzadd(const Fq& a, const Fq& b) {
        Fq r;
#pragma unroll
        for (int i = 0; i < 9; i++) r.v[i] = a.v[i] + b.v[i];
        return r;
    }

    static __device__ __forceinline__ Fq2 triple2(const Fq2& a) {
        return Fq2{red8(A::fq_add(A::fq_add(a.c0, a.c0), a.c0)),
                   red8(A::fq_add(A::fq_add(a.c1, a.c1), a.c1))};
    }

    // ---- XYZZ points over Fp2; ZZ = 0 is the identity
    static __device__ __forceinline__ QXyzz2 identity() {
        return QXyzz2{zero2(), zero2(), zero2(), zero2()};
    }

    static __device__ __forceinline__ bool is_id(const QXyzz2& p) { return is_zero2(p.ZZ); }

    static __device__ __forceinline__ QXyzz2 select(bool c, const QXyzz2& a, const QXyzz2& b) {
        return QXyzz2{sel2(c, a.X, b.X), sel2(c, a.Y, b.Y), sel2(c, a.ZZ, b.ZZ), sel2(c, a.ZZZ, b.ZZZ)};
    }

    static __device__ __forceinline__ void dbl_core(const Fq2& X, const Fq2& Y, Fq2& X3, Fq2& Y3,
                                                    Fq2& V, Fq2& W) {
        const Fq2 U = dbl2(Y);
        V = sqr2(U);
        W = mul2(U, V);
        const Fq2 S = mul2(X, V);
        const Fq2 M = triple2(sqr2(X));
        X3 = sub2(sub2(sqr2(M), S), S);
        Y3 = sub2(mul2(M, sub2(S, X3)), mul2(W, Y));
    }

    // 2 (x, y) for an affine point (y = 0 gives ZZ = 0, the identity)
    static __device__ __forceinline__ QXyzz2 dbl_aff(const Fq2& x, const Fq2& y) {
        QXyzz2 r;
        dbl_core(x, y, r.X, r.Y, r.ZZ, r.ZZZ);
        return r;
    }

    // 2 P; the identity stays the identity (ZZ3 = V ZZ = 0)
    static __device__ __forceinline__ QXyzz2 dbl(const QXyzz2& p) {
        QXyzz2 r;
        Fq2 V, W;
        dbl_core(p.X, p.Y, r.X, r.Y, V, W);
        r.ZZ = mul2(V, p.ZZ);
        r.ZZZ = mul2(W, p.ZZZ);
        return r;
    }

    // the value an accumulator holds at a bucket run's start: never used (the
    // run's first addition selects the point), ZZ = ZZZ = 1
    static __device__ __forceinline__ QXyzz2 run_start() {
        return QXyzz2{zero2(), zero2(), one2(), one2()};
    }

    // acc += (x2, y2) in the bucket-accumulation loop (x2, y2 < 2p, never the
    // identity); acc_id says acc is the identity (a run start, or a
    // cancellation earlier in the run).  Doubling and cancellation have P = 0;
    // the exact tests and the doubling run in a branch a wave takes only when
    // one of its lanes needs it (a ballot makes the branch wave-uniform, so
    // the compiler keeps it a branch instead of computing the doubling on
    // every step and selecting).
    // Products: mul2m (merged reductions); bounds (multiples of p): acc < 2p
    // per component, P, R < 3.1p, PP < 1.3p, PPP, Q < 1.1p, T = 2Q + PPP < 4p,
    // X3 < 5.3p and Y3 < 3.2p before their conditional subtractions, every
    // stored component < 2p after them.
    static __device__ __forceinline__ void madd_run(QXyzz2& acc, bool& acc_id, const Fq2& x2,
                                                    const Fq2& y2) {
        const Fq2 P = mul2m<2>(x2, acc.ZZ, acc.X);   // U2 - X1 + 2p
        const Fq2 R = mul2m<2>(y2, acc.ZZZ, acc.Y);  // S2 - Y1 + 2p
        // PP = (P0 + P1)(P0 - P1 + 4p) + 2 P0 P1 u
        Fq2 PP;
        A::fq_mul_sub_x2<0, 1, 0, 1>(A::fq_add(P.c0, P.c1), A::fq_sub_lazy<4>(P.c0, P.c1), P.c0, PP.c0,
                                     lzadd(P.c0, P.c0), P.c1, P.c0, PP.c1);
        const Fq2 PPP = mul2m<0>(P, PP, P);
        const Fq2 Q = mul2m<0>(acc.X, PP, P);
        Fq2 T;  // 2Q + PPP, limbs sums of three normalised limbs
#pragma unroll
        for (int i = 0; i < 9; i++) {
            T.c0.v[i] = Q.c0.v[i] + Q.c0.v[i] + PPP.c0.v[i];
            T.c1.v[i] = Q.c1.v[i] + Q.c1.v[i] + PPP.c1.v[i];
        }
        QXyzz2 r;
        // X3 = R^2 - T: (R0 + R1)(R0 - R1 + 4p) - T0 + 4p, 2 R0 R1 - T1 + 4p
        A::fq_mul_sub_x2<4, 3, 4, 3>(A::fq_add(R.c0, R.c1), A::fq_sub_lazy<4>(R.c0, R.c1), T.c0, r.X.c0,
                                     lzadd(R.c0, R.c0), R.c1, T.c1, r.X.c1);
        r.X = Fq2{red8(r.X.c0), red8(r.X.c1)};
        // Y3 = R (Q - X3) - Y PPP
        const Fq2 M2 = mul2m<0>(acc.Y, PPP, P);
        const Fq2 TT{A::fq_sub<2>(Q.c0, r.X.c0), A::fq_sub<2>(Q.c1, r.X.c1)};  // < 3.1p
        r.Y = mul2m<2, 4>(R, TT, M2);
        r.Y = Fq2{red4(r.Y.c0), red4(r.Y.c1)};
        r.ZZ = mul2m<0>(acc.ZZ, PP, P);
        r.ZZZ = mul2m<0>(acc.ZZZ, PPP, P);
        bool cancel = false;
        const bool ex = !acc_id && is_zero2(PP);  // P = 0 iff PP = 0
        if (__builtin_expect(__ballot(ex) != 0, 0)) {
            const bool d = ex && is_zero2(Fq2{red4(R.c0), red4(R.c1)});
            cancel = ex && !d;
            const QXyzz2 dd = dbl_aff(x2, y2);
            r = select(d, dd, r);
        }
        const QXyzz2 q{x2, y2, one2(), one2()};
        acc = select(acc_id, q, r);
        acc_id = cancel;
    }

    // acc += q (both XYZZ, components < 2p), on the merged-reduction products
    // of madd_run with the same bounds (U1, S1 < 1.1p)
    static __device__ __forceinline__ void add(QXyzz2& acc, const QXyzz2& q) {
        const bool p_id = is_id(acc);
        const bool q_id = is_id(q);
        const Fq2 U1 = mul2m<0>(acc.X, q.ZZ, acc.X);
        const Fq2 S1 = mul2m<0>(acc.Y, q.ZZZ, acc.X);
        const Fq2 P = mul2m<2>(q.X, acc.ZZ, U1);    // U2 - U1 + 2p
        const Fq2 R = mul2m<2>(q.Y, acc.ZZZ, S1);  // S2 - S1 + 2p
        Fq2 PP;
        A::fq_mul_sub_x2<0, 1, 0, 1>(A::fq_add(P.c0, P.c1), A::fq_sub_lazy<4>(P.c0, P.c1), P.c0, PP.c0,
                                     lzadd(P.c0, P.c0), P.c1, P.c0, PP.c1);
        const Fq2 PPP = mul2m<0>(P, PP, P);
        const Fq2 Q = mul2m<0>(U1, PP, P);
        Fq2 T;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            T.c0.v[i] = Q.c0.v[i] + Q.c0.v[i] + PPP.c0.v[i];
            T.c1.v[i] = Q.c1.v[i] + Q.c1.v[i] + PPP.c1.v[i];
        }
        QXyzz2 r;
        A::fq_mul_sub_x2<4, 3, 4, 3>(A::fq_add(R.c0, R.c1), A::fq_sub_lazy<4>(R.c0, R.c1), T.c0, r.X.c0,
                                     lzadd(R.c0, R.c0), R.c1, T.c1, r.X.c1);
        r.X = Fq2{red8(r.X.c0), red8(r.X.c1)};
        const Fq2 M2 = mul2m<0>(S1, PPP, P);
        const Fq2 TT{A::fq_sub<2>(Q.c0, r.X.c0), A::fq_sub<2>(Q.c1, r.X.c1)};
        r.Y = mul2m<2, 4>(R, TT, M2);
        r.Y = Fq2{red4(r.Y.c0), red4(r.Y.c1)};
        r.ZZ = mul2m<0>(mul2m<0>(acc.ZZ, q.ZZ, P), PP, P);
        r.ZZZ = mul2m<0>(mul2m<0>(acc.ZZZ, q.ZZZ, P), PPP, P);
        const bool d = is_zero2(PP) && is_zero2(Fq2{red4(R.c0), red4(R.c1)}) && !p_id && !q_id;
        if (__builtin_expect(d, 0)) r = dbl(acc);
        r = select(p_id, q, r);
        acc = select(q_id, acc, r);
    }

    // acc += (x2, y2) with q_id: the affine point is the identity
    static __device__ __forceinline__ void madd(QXyzz2& acc, const Fq2& x2, const Fq2& y2, bool q_id) {
        const QXyzz2 q{x2, y2, one2(), one2()};
        add(acc, select(q_id, identity(), q));
    }

    // ---- conversions from / to the 2^256 Montgomery form of 32-bit limbs
    static __device__ __forceinline__ Fq2 from_fp2(const Fp& c0, const Fp& c1) {
        return Fq2{A::fq_from_fp_canon(canon32<Bn254Fp>(c0)), A::fq_from_fp_canon(canon32<Bn254Fp>(c1))};
    }

    // a < 2p -> canonical 32-bit form
    static __device__ __forceinline__ Fp to_fp(const Fq& a) { return A::fq_to_fp(a); }

    // Fermat inverse of an Fp2 element: conj(a) / (a0^2 + a1^2); 0 -> 0
    static __device__ Fq2 inv2(const Fq2& a) {
        const Fq n = red4(A::fq_add(A::fq_sqr(a.c0), A::fq_sqr(a.c1)));
        const Fq ni = A::fq_inv(n);
        return Fq2{A::fq_mul(a.c0, ni), red4(A::fq_sub<2>(fq_zero(), A::fq_mul(a.c1, ni)))};
    }
};

// ---- global memory: table slots 2e (c0 parts) and 2e + 1 (c1 parts)
__device__ __forceinline__ QAff2 load_qtab2(const uint4* base, uint64_t e, bool neg) {
    const QAff a = load_qtab(base, 2 * e, neg);
    const QAff b = load_qtab(base, 2 * e + 1, neg);
    return QAff2{Fq2{a.x, b.x}, Fq2{a.y, b.y}};
}

__device__ __forceinline__ void store_qtab2(uint4* base, uint64_t e, const QAff2& p) {
    G2Arith::A::store_qtab(base, 2 * e, QAff{p.x.c0, p.y.c0});
    G2Arith::A::store_qtab(base, 2 * e + 1, QAff{p.x.c1, p.y.c1});
}

__device__ __forceinline__ QXyzz2 load_qxyzz2(const uint4* base, uint64_t i) {
    const QXyzz a = load_qxyzz(base, 2 * i);
    const QXyzz b = load_qxyzz(base, 2 * i + 1);
    return QXyzz2{Fq2{a.X, b.X}, Fq2{a.Y, b.Y}, Fq2{a.ZZ, b.ZZ}, Fq2{a.ZZZ, b.ZZZ}};
}

__device__ __forceinline__ void store_qxyzz2(uint4* base, uint64_t i, const QXyzz2& p) {
    store_qxyzz(base, 2 * i, QXyzz{p.X.c0, p.Y.c0, p.ZZ.c0, p.ZZZ.c0});
    store_qxyzz(base, 2 * i + 1, QXyzz{p.X.c1, p.Y.c1, p.ZZ.c1, p.ZZZ.c1});
}

// host form: two 128-byte XYZZ records (c0 parts, then c1 parts), canonical
// 2^256 Montgomery form
__device__ __forceinline__ void store_host_xyzz2(uint4* base, uint64_t i, const QXyzz2& p) {
    using G = G2Arith;
    Xyzz a, b;
    a.X = G::to_fp(p.X.c0); a.Y = G::to_fp(p.Y.c0); a.ZZ = G::to_fp(p.ZZ.c0); a.ZZZ = G::to_fp(p.ZZZ.c0);
    b.X = G::to_fp(p.X.c1); b.Y = G::to_fp(p.Y.c1); b.ZZ = G::to_fp(p.ZZ.c1); b.ZZZ = G::to_fp(p.ZZZ.c1);
    store_xyzz(base, 2 * i, a);
    store_xyzz(base, 2 * i + 1, b);
}

__device__ __forceinline__ QXyzz2 load_host_xyzz2(const uint4* base, uint64_t i) {
    const Xyzz a = load_xyzz(base, 2 * i), b = load_xyzz(base, 2 * i + 1);
    using G = G2Arith;
    return QXyzz2{G::from_fp2(a.X, b.X), G::from_fp2(a.Y, b.Y), G::from_fp2(a.ZZ, b.ZZ),
                  G::from_fp2(a.ZZZ, b.ZZZ)};
}

}  // namespace zk
