// HIP kernels of the BN254 G1 MSM for gfx950 (wave64).
//
// Pipeline (per batch of MSMs, all on one stream):
//   k_build_table     bases -> [P, phi(P)] table              (glv_kernels.metal:260-297)
//   k_decompose       scalars -> GLV halves -> signed digits  (glv_kernels.metal:172-256,
//                                                              msm_kernels.metal:399-439)
//   k_sort_hist / k_scan_* / k_sort_scatter, two passes
//                     stable LSD radix sort of bucket keys per segment
//                     (replaces the CPU counting sort, MSMEngine.swift:595-661)
//   k_accumulate      load-balanced XYZZ mixed additions into buckets
//                     (replaces msm_kernels.metal:16-62)
//   k_merge_runs / k_merge_long   bucket partials that crossed lane boundaries
//   k_br_rowcol / k_br_bits  bucket reduction: window sum = U + sum_b 2^b Q_b,
//                     Q_b bit-plane sums of the bucket index (replaces the
//                     running sums of msm_kernels.metal:152-248)
// The weighting by 2^b, the window combine (Horner over windows) and the
// affine normalisation are a short serial tail that runs on the host
// (MSMEngine.swift:899-904 does its Horner there too).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "curve29.hpp"
#include "curve_g2.hpp"
#include "glv.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace zk {

// BN254 G1 over Fp: the GLV curve of the MSM, the SRS import / generation and
// the primitive tests; Grumpkin over Fr: the sibling curve (no GLV)
using Ap = Arith<Bn254Fp>;
using Ar = Arith<Bn254Fr>;

// ------------------------------------------------------------------ helpers

// largest j in [0, n) with pre[j] <= x (pre ascending, pre[0] == 0)
__device__ __forceinline__ uint32_t upper_index(const uint64_t* pre, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n;  // invariant: pre[lo] <= x < pre[hi] (pre[n] = inf)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Sort key of a bucket index x < 2^15: low 7 bits in byte 0, the rest in
// byte 1.  A real key's low byte is then <= 0x7f, so the first radix pass
// puts every "no bucket" entry (0xFFFF) in bin 0xFF, alone, and can drop them.
__device__ __forceinline__ uint32_t key_encode(uint32_t x) { return (x & 0x7fu) | ((x >> 7) << 8); }
__device__ __forceinline__ uint32_t key_decode(uint32_t k) { return (k & 0x7fu) | ((k >> 8) << 7); }

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Block of NW waves: exclusive prefix of v, block total in *total (lds: NW words)
template <int NW = 4>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds,
                                                         uint32_t* total) {
    const uint32_t inc = wave_inclusive_scan(v);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 63u) lds[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)NW; i++) {
        uint32_t x = lds[i];
        off += (i < w) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// Sum over aligned groups of P lanes (P a power of two <= 64, uniform across
// the wave); the group's first lane returns the group total.  Cross-lane data
// moves with ds_bpermute shuffles, no LDS, no barriers.
__device__ __forceinline__ Fq shfl_down_fq(const Fq& a, uint32_t d) {
    Fq r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = __shfl_down(a.v[i], d);
    return r;
}

__device__ __forceinline__ QXyzz shfl_down_qxyzz(const QXyzz& a, uint32_t d) {
    QXyzz r;
    r.X = shfl_down_fq(a.X, d);
    r.Y = shfl_down_fq(a.Y, d);
    r.ZZ = shfl_down_fq(a.ZZ, d);
    r.ZZZ = shfl_down_fq(a.ZZZ, d);
    return r;
}

// Curve policies of the bucket kernels (accumulation, merge, reduction): the
// stored point type and its loads / stores, the affine table point, and the
// additions.  G1Curve<F>: XYZZ over one field (BN254 G1 over Fp, Grumpkin
// over Fr), one 128-byte table slot per entry, 144-byte points.  G2Curve:
// XYZZ over Fp2 (BN254 G2), two slots per entry, two 144-byte halves per
// point.  kSlots: table slots per entry.
template <class F>
struct G1Curve {
    using Pt = QXyzz;
    using Af = QAff;
    static constexpr uint32_t kSlots = 1;
    static constexpr uint64_t kPtBytes = kBucketBytes;
    static __device__ __forceinline__ Pt identity() { return qxyzz_identity(); }
    static __device__ __forceinline__ Pt select(bool c, const Pt& a, const Pt& b) { return qxyzz_select(c, a, b); }
    static __device__ __forceinline__ Pt run_start() { return Arith<F>::qxyzz_run_start(); }
    static __device__ __forceinline__ void madd_run(Pt& a, bool& id, const Af& q) {
        Arith<F>::qxyzz_madd_run(a, id, q.x, q.y);
    }
    static __device__ __forceinline__ void add(Pt& a, const Pt& b) { Arith<F>::qxyzz_add(a, b); }
    static __device__ __forceinline__ Pt shfl_down(const Pt& a, uint32_t d) { return shfl_down_qxyzz(a, d); }
    static __device__ __forceinline__ Af load_tab(const uint4* t, uint64_t e, bool neg) { return load_qtab(t, e, neg); }
    static __device__ __forceinline__ Pt load(const uint4* b, uint64_t i) { return load_qxyzz(b, i); }
    static __device__ __forceinline__ void store(uint4* b, uint64_t i, const Pt& p) { store_qxyzz(b, i, p); }
    static __device__ __forceinline__ void store_host(uint4* b, uint64_t i, const Pt& p) {
        store_xyzz(b, i, Arith<F>::qxyzz_to_xyzz(p));
    }
    static __device__ __forceinline__ Pt dbl(const Pt& p) { return Arith<F>::qxyzz_dbl(p); }
    static __device__ __forceinline__ Pt dbl_k(const Pt& p, uint32_t k) { return Arith<F>::qxyzz_dbl_k(p, k); }
    // result slot i (96 B): the normalised Jacobian X | Y | Z in the 2^256
    // Montgomery form, canonical: (x, y, 1), the identity (1, 1, 0)
    static __device__ __forceinline__ void store_result(uint4* b, uint64_t i, const Pt& p) {
        using A = Arith<F>;
        const Fp one = A::fq_to_fp(A::fq_one());
        Fp x = one, y = one, z = fp_zero();
        if (!A::qxyzz_is_id(p)) {
            // 1/ZZ = ZZZ / (ZZ ZZZ), 1/ZZZ = ZZ / (ZZ ZZZ): one inversion
            const Fq inv = A::fq_inv(A::fq_mul(p.ZZ, p.ZZZ));
            x = A::fq_to_fp(A::fq_mul(p.X, A::fq_mul(p.ZZZ, inv)));
            y = A::fq_to_fp(A::fq_mul(p.Y, A::fq_mul(p.ZZ, inv)));
            z = one;
        }
        uint4* o = b + i * 6;
        const Fp* f[3] = {&x, &y, &z};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            o[2 * k] = make_uint4(f[k]->v[0], f[k]->v[1], f[k]->v[2], f[k]->v[3]);
            o[2 * k + 1] = make_uint4(f[k]->v[4], f[k]->v[5], f[k]->v[6], f[k]->v[7]);
        }
    }
};

struct G2Curve {
    using Pt = QXyzz2;
    using Af = QAff2;
    static constexpr uint32_t kSlots = 2;
    static constexpr uint64_t kPtBytes = 2 * kBucketBytes;
    static __device__ __forceinline__ Pt identity() { return G2Arith::identity(); }
    static __device__ __forceinline__ Pt select(bool c, const Pt& a, const Pt& b) { return G2Arith::select(c, a, b); }
    static __device__ __forceinline__ Pt run_start() { return G2Arith::run_start(); }
    static __device__ __forceinline__ void madd_run(Pt& a, bool& id, const Af& q) {
        G2Arith::madd_run(a, id, q.x, q.y);
    }
    static __device__ __forceinline__ void add(Pt& a, const Pt& b) { G2Arith::add(a, b); }
    static __device__ __forceinline__ Pt shfl_down(const Pt& a, uint32_t d) {
        const QXyzz x = shfl_down_qxyzz(QXyzz{a.X.c0, a.Y.c0, a.ZZ.c0, a.ZZZ.c0}, d);
        const QXyzz y = shfl_down_qxyzz(QXyzz{a.X.c1, a.Y.c1, a.ZZ.c1, a.ZZZ.c1}, d);
        return Pt{Fq2{x.X, y.X}, Fq2{x.Y, y.Y}, Fq2{x.ZZ, y.ZZ}, Fq2{x.ZZZ, y.ZZZ}};
    }
    static __device__ __forceinline__ Af load_tab(const uint4* t, uint64_t e, bool neg) { return load_qtab2(t, e, neg); }
    static __device__ __forceinline__ Pt load(const uint4* b, uint64_t i) { return load_qxyzz2(b, i); }
    static __device__ __forceinline__ void store(uint4* b, uint64_t i, const Pt& p) { store_qxyzz2(b, i, p); }
    static __device__ __forceinline__ void store_host(uint4* b, uint64_t i, const Pt& p) { store_host_xyzz2(b, i, p); }
};

// Lane-sum then group tree with ONE addition site: steps 0..nload-1 add the
// lane's own elements (load(e) returns the e-th, or the identity), the next
// log2(P) steps add the partner's partial (shuffle tree over aligned groups of
// P lanes).  nload and P must be wave-uniform.  Keeping a single inlined
// addition halves the code of the latency-bound reduction kernels.
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d));
    return v;
}

// point i of b where `valid`, else the identity: limb-wise selects (a
// conditional load of a whole point as one value made the compiler stage it
// through scratch memory); an invalid lane reads point 0, which every such
// buffer holds
template <class C>
__device__ __forceinline__ typename C::Pt load_or_id(const uint4* b, uint64_t i, bool valid) {
    return C::select(valid, C::load(b, valid ? i : 0), C::identity());
}

template <class C, class Load>
__device__ __forceinline__ typename C::Pt sum_then_tree(uint32_t nload, uint32_t P, Load load) {
    const uint32_t lane = __lane_id();
    uint32_t levels = 0;
    while ((1u << levels) < P) levels++;
    typename C::Pt acc = C::identity();
#pragma unroll 1
    for (uint32_t st = 0; st < nload + levels; st++) {
        typename C::Pt q;
        if (st < nload) {
            q = load(st);
        } else {
            const uint32_t s = P >> (st - nload + 1);
            q = C::select((lane & (P - 1)) >= s, C::identity(), C::shfl_down(acc, s));
        }
        C::add(acc, q);
    }
    return acc;
}

// -------------------------------------------------------- base table / phi
// table[2i] = P_i, table[2i+1] = phi(P_i) = (beta x, y): affine, coordinates
// in the 2^261 Montgomery form, canonical, as 29-bit limbs with p - y beside y
// (store_qtab, kTableEntryBytes per entry).  (0,0) maps to (0,0).  After the
// n pairs comes a bitmap, bit i set when P_i is the identity: the
// decomposition gives such a point no digits, so the accumulation never adds
// an identity base.

// one 64-bit bitmap word per wave; the wave covers points [64k, 64k + 64)
// (256-thread blocks over consecutive points) and lane 0 is active whenever
// any lane is, since lanes past n leave before this call
__device__ __forceinline__ void store_id_bit(uint64_t* idmask, uint64_t i, bool is_id) {
    const uint64_t m = __ballot(is_id);
    if (__lane_id() == 0) idmask[i >> 6] = m;
}

__device__ __forceinline__ void store_table_pair(uint4* table, uint64_t* idmask, uint64_t i,
                                                 const Aff& a) {
    QAff q;
    q.x = Ap::fq_from_fp_canon(a.x);
    q.y = Ap::fq_from_fp_canon(a.y);
    store_id_bit(idmask, i, qaff_is_id(q));
    Ap::store_qtab(table, 2 * i, q);
    q.x = Ap::fq_csub<1>(Ap::fq_mul(Ap::fq_beta(), q.x));
    Ap::store_qtab(table, 2 * i + 1, q);
}

// BN254 G2 table: point i (x.c0, x.c1, y.c0, y.c1 in the 2^256 form, 128 B)
// -> slots 2i (the c0 parts) and 2i + 1 (the c1 parts), canonical; an
// all-zero point is the identity (bitmap bit).
__global__ __launch_bounds__(256) void k_build_table_g2(const uint4* __restrict__ pts,
                                                        uint4* __restrict__ table,
                                                        uint64_t* __restrict__ idmask, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4* p = pts + i * 8;
    const QAff2 q{G2Arith::from_fp2(load_fp(p), load_fp(p + 2)), G2Arith::from_fp2(load_fp(p + 4), load_fp(p + 6))};
    store_id_bit(idmask, i, qaff_is_id(QAff{q.x.c0, q.y.c0}) && qaff_is_id(QAff{q.x.c1, q.y.c1}));
    store_qtab2(table, i, q);
}

// SRS import from an SRS file's G1 section (Sources/zkMetal/KZG/
// SRSFileFormat.swift:92-160, TrustedSetup.swift:58-70,686-696): point i is
// x then y, 32-byte big-endian standard-form integers; (0,0) is the identity.
// Each point is checked canonical (x, y < p) and on the curve y^2 = x^3 + 3;
// `bad` counts the points that are not.  Writes the table pair of every point.
__device__ __forceinline__ Fp fp_from_be(const uint8_t* b) {
    Fp r;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint8_t* w = b + 28 - 4 * j;
        r.v[j] = ((uint32_t)w[0] << 24) | ((uint32_t)w[1] << 16) | ((uint32_t)w[2] << 8) | w[3];
    }
    return r;
}

__device__ __forceinline__ bool fp_lt_p(const Fp& a) {
    unsigned br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) (void)__builtin_subc(a.v[i], p_limb(i), br, &br);
    return br != 0;
}

__device__ __forceinline__ bool fq_eq(const Fq& a, const Fq& b) {
    uint32_t d = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) d |= a.v[i] ^ b.v[i];
    return d == 0;
}

__global__ __launch_bounds__(256) void k_import_be(const uint8_t* __restrict__ raw, uint64_t n,
                                                   uint4* __restrict__ table,
                                                   uint64_t* __restrict__ idmask,
                                                   uint32_t* __restrict__ bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Fp x = fp_from_be(raw + 64 * i), y = fp_from_be(raw + 64 * i + 32);
    Fq k2;  // 2^522 mod p: a Montgomery product with it maps x to x * 2^261
    k2.v[0] = 0x059bac10u; k2.v[1] = 0x0d1503a3u; k2.v[2] = 0x018016b8u; k2.v[3] = 0x10ab0ca8u;
    k2.v[4] = 0x02632639u; k2.v[5] = 0x02c0169fu; k2.v[6] = 0x169bfd53u; k2.v[7] = 0x11869d4cu;
    k2.v[8] = 0x002a11a6u;
    QAff q;
    q.x = Ap::fq_csub<1>(Ap::fq_mul(fq_slice(x), k2));
    q.y = Ap::fq_csub<1>(Ap::fq_mul(fq_slice(y), k2));
    bool ok = fp_lt_p(x) && fp_lt_p(y);
    if (!qaff_is_id(q)) {
        Fq b3;  // 3 * 2^261 mod p
        b3.v[0] = 0x00766463u; b3.v[1] = 0x1c54760au; b3.v[2] = 0x08f6927au; b3.v[3] = 0x03e40c4du;
        b3.v[4] = 0x1fea4f2bu; b3.v[5] = 0x17c6c26au; b3.v[6] = 0x157fe417u; b3.v[7] = 0x0f8056f9u;
        b3.v[8] = 0x002958a2u;
        const Fq lhs = Ap::fq_csub<1>(Ap::fq_mul(q.y, q.y));                                // < p
        const Fq rhs = Ap::fq_csub<1>(Ap::fq_csub<2>(Ap::fq_add(Ap::fq_mul(Ap::fq_mul(q.x, q.x), q.x), b3)));  // < 2.5p -> < p
        ok = ok && fq_eq(lhs, rhs);
    }
    if (!ok) atomicAdd(bad, 1u);
    store_id_bit(idmask, i, qaff_is_id(q));
    Ap::store_qtab(table, 2 * i, q);
    q.x = Ap::fq_csub<1>(Ap::fq_mul(Ap::fq_beta(), q.x));
    Ap::store_qtab(table, 2 * i + 1, q);
}

// table entry EPS i (the base itself) back to the 2^256 Montgomery form
// (64 B per point); EPS = table entries per base (2: P and phi(P), 1: P)
template <class F, int EPS>
__global__ __launch_bounds__(256) void k_table_export(const uint4* __restrict__ table, uint64_t n,
                                                      uint4* __restrict__ pts) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    QAff q = load_qtab(table, EPS * i, false);
    Aff a;
    a.x = Arith<F>::fq_to_fp(q.x);
    a.y = Arith<F>::fq_to_fp(q.y);
    store_aff(pts, i, a);
}

// The block's 256 table rows (EPS entries per point: P, and phi(P) for the
// GLV curve) are staged in LDS (a row per point, padded by 16 B) and written
// out as consecutive 16-byte pieces across the block, so every store
// instruction of a wave covers 1 KB of contiguous table.
template <class F, int EPS>
__global__ __launch_bounds__(256) void k_build_table(const uint4* __restrict__ pts,
                                                     uint4* __restrict__ table,
                                                     uint64_t* __restrict__ idmask, uint64_t n) {
    using A = Arith<F>;
    constexpr uint32_t kPair = EPS * kTableEntryBytes / 16;  // uint4 per point
    constexpr uint32_t kRow = kPair + 1;
    __shared__ uint4 stage[256 * kRow];
    const uint64_t i0 = (uint64_t)blockIdx.x * 256;
    const uint64_t i = i0 + threadIdx.x;
    if (i < n) {
        Aff a = load_aff(pts, i);
        QAff q;
        q.x = A::fq_from_fp_canon(canon32<F>(a.x));
        q.y = A::fq_from_fp_canon(canon32<F>(a.y));
        store_id_bit(idmask, i, qaff_is_id(q));
        uint4* row = stage + threadIdx.x * kRow;
        A::store_qtab(row, 0, q);
        if constexpr (EPS == 2) {
            q.x = A::template fq_csub<1>(A::fq_mul(A::fq_beta(), q.x));
            A::store_qtab(row, 1, q);
        }
    }
    __syncthreads();
    const uint32_t words = (uint32_t)min<uint64_t>(256, n - i0) * kPair;
    uint4* dst = table + i0 * kPair;
    for (uint32_t w = threadIdx.x; w < words; w += 256) dst[w] = stage[(w / kPair) * kRow + w % kPair];
}

// Window-shifted copies of a BN254 G1 base table (the resident SRS): for
// windows w = 1 .. nwin-1, entries w * stride + 2i and w * stride + 2i + 1
// hold 2^(c w) P_i and phi(2^(c w) P_i) = (beta x, y), affine and canonical
// like the table's own entries (window 0).  A digit of window w then adds
// its base's shifted copy, so every window of an MSM can share one set of
// buckets and the window sums need no doublings to combine.  One lane per
// base: c Jacobian doublings per window (dbl-2009-l, exact for points of odd
// order), one inversion back to affine; identity bases are skipped (they
// carry no digits).
__global__ __launch_bounds__(256) void k_shift_windows(uint4* __restrict__ table, uint64_t n,
                                                       uint64_t stride, uint32_t c, uint32_t nwin) {
    using A = Arith<Bn254Fp>;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t* idmask = reinterpret_cast<const uint64_t*>(table + n * 2 * (kTableEntryBytes / 16));
    if ((idmask[i >> 6] >> (i & 63)) & 1) return;
    QAff q = load_qtab(table, 2 * i, false);
    Fq X = q.x, Y = q.y, Z = A::fq_one();  // < p each
#pragma unroll 1
    for (uint32_t w = 1; w < nwin; w++) {
#pragma unroll 1
        for (uint32_t k = 0; k < c; k++) A::qjac_dbl(X, Y, Z);  // X, Y, Z < 2p
        const Fq zi = A::fq_inv(Z);                              // < 1.5p
        const Fq zi2 = A::fq_sqr(zi);
        q.x = A::template fq_csub<1>(A::fq_mul(X, zi2));             // X / Z^2, < p
        q.y = A::template fq_csub<1>(A::fq_mul(Y, A::fq_mul(zi2, zi)));  // Y / Z^3, < p
        A::store_qtab(table, w * stride + 2 * i, q);
        QAff ph = q;
        ph.x = A::template fq_csub<1>(A::fq_mul(A::fq_beta(), q.x));
        A::store_qtab(table, w * stride + 2 * i + 1, ph);
        X = q.x;
        Y = q.y;
        Z = A::fq_one();
    }
}

// --------------------------------------------------- GLV + signed digits

// Grumpkin's group order is the BN254 base-field prime p: scalars are reduced
// mod p (GrumpkinMSMEngine.swift:178-196) and recoded over 255 bits, no GLV.
__device__ __forceinline__ void reduce_mod_p(uint32_t k[8]) {
#pragma unroll 1
    for (int it = 0; it < 6; it++) {
        uint32_t d[8];
        unsigned br = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) d[i] = __builtin_subc(k[i], p_limb(i), br, &br);
        if (br) break;
#pragma unroll
        for (int i = 0; i < 8; i++) k[i] = d[i];
    }
}

// BN254 G2's group order is r: scalars are reduced mod r
// (BN254G2MSMEngine.swift:214-232) and recoded like Grumpkin's.
__device__ __forceinline__ void reduce_mod_r(uint32_t k[8]) {
#pragma unroll 1
    for (int it = 0; it < 6; it++) {
        uint32_t d[8];
        unsigned br = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) d[i] = __builtin_subc(k[i], r_limb(i), br, &br);
        if (br) break;
#pragma unroll
        for (int i = 0; i < 8; i++) k[i] = d[i];
    }
}

// c-bit window w of a 256-bit value, from 64-bit words by shifts (w is the
// same in every lane, so the word choice is a uniform branch, no indexing)
__device__ __forceinline__ uint32_t window_bits256(const uint32_t k[8], uint32_t w, uint32_t c) {
    const uint32_t off = w * c, q = off >> 6, sh = off & 63u;
    const uint64_t w0 = ((uint64_t)k[1] << 32) | k[0], w1 = ((uint64_t)k[3] << 32) | k[2];
    const uint64_t w2 = ((uint64_t)k[5] << 32) | k[4], w3 = ((uint64_t)k[7] << 32) | k[6];
    uint64_t lo = 0, hi = 0;
    if (q == 0) { lo = w0; hi = w1; }
    else if (q == 1) { lo = w1; hi = w2; }
    else if (q == 2) { lo = w2; hi = w3; }
    else if (q == 3) { lo = w3; }
    const uint64_t x = (lo >> sh) | (sh ? hi << (64u - sh) : 0ull);
    return (uint32_t)x & ((1u << c) - 1u);
}

// Block b decomposes the kSortTile / EPS scalars of first-pass sort tile tin
// of MSM j (msm_dtile0[j] <= b < msm_dtile0[j+1], tin = b - msm_dtile0[j]):
// their entries are tile tin of every window's entries, so the block also
// counts the tile's 256-bin histogram of key byte 0 for each window (LDS
// atomics) and writes it where the scan expects it; the first pass then
// needs no histogram kernel.  EPS = 2: BN254 G1, GLV halves k1 (entry 2i,
// base P_i) and k2 (entry 2i + 1, phi(P_i)); EPS = 1: one entry per scalar
// over 255 bits, the scalar reduced mod r (kModR, BN254 G2) or mod p
// (Grumpkin).  Window w's entries: segment seg0 + w / wps, at local window
// v = w % wps, i.e. tiles v * ntw + tin and positions v * ntw * kSortTile +
// (entry) of the segment, ntw = the window's tiles; a window-shifted MSM
// (wps > 1) pads each window to whole tiles with "no bucket" digits.
constexpr int kDecThreads = 256;
static_assert(kSortTile <= (1 << 15) && kSortTile / 2 / kDecThreads >= 1, "decomposition tiling");

template <int EPS, bool kModR>
__global__ __launch_bounds__(kDecThreads) void k_decompose(const MsmDesc* __restrict__ msms,
                                                          const uint64_t* __restrict__ msm_dtile0,
                                                          uint32_t nmsm,
                                                          const SegDesc* __restrict__ segs,
                                                          const uint4* __restrict__ scalars,
                                                          const uint64_t* __restrict__ idmask,
                                                          uint2* __restrict__ digits,
                                                          uint32_t* __restrict__ hist,
                                                          uint32_t* __restrict__ err) {
    // windows (nwin <= 64) x key byte 0, two 16-bit counts per word (a tile
    // has at most kSortTile <= 2^15 entries, so no count reaches 2^16)
    __shared__ uint32_t cnt[kMaxWindows * 128];
    const uint32_t b = blockIdx.x;
    const uint32_t j = upper_index(msm_dtile0, nmsm, b);
    const MsmDesc m = msms[j];
    const uint32_t tin = b - (uint32_t)msm_dtile0[j];
    for (uint32_t x = threadIdx.x; x < m.nwin * 128u; x += kDecThreads) cnt[x] = 0;
    __syncthreads();
    const uint32_t c = m.c, half = 1u << (c - 1), full = 1u << c;
    bool carry_left = false;
    constexpr int kPer = kSortTile / EPS / kDecThreads;  // scalars per thread
    const uint32_t ntw = (EPS * m.n + kSortTile - 1) / kSortTile;  // tiles per window
    const uint64_t vstride = (uint64_t)ntw * kSortTile;             // positions per local window
    for (int r = 0; r < kPer; r++) {
        const uint64_t i = (uint64_t)tin * (kSortTile / EPS) + (uint64_t)r * kDecThreads + threadIdx.x;
        const bool live = i < m.n;
        // past the MSM's scalars: done, unless its windows are padded to whole
        // tiles (window-shifted), whose tails take "no bucket" digits
        if (!live && m.wps == 1) break;
        uint32_t k[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (live) {
            const uint64_t g = m.scalar0 + i;
            uint4 s0 = scalars[2 * g], s1 = scalars[2 * g + 1];
            k[0] = s0.x; k[1] = s0.y; k[2] = s0.z; k[3] = s0.w;
            k[4] = s1.x; k[5] = s1.y; k[6] = s1.z; k[7] = s1.w;
            if (m.form == 1) fr_from_mont(k);
            const uint64_t pt = m.table0 / EPS + i;  // the base's point index in the table
            if ((idmask[pt >> 6] >> (pt & 63)) & 1) {
#pragma unroll
                for (int q = 0; q < 8; q++) k[q] = 0;  // identity base: no digits
            }
        }
        if constexpr (EPS == 1) {
            if constexpr (kModR) reduce_mod_r(k);
            else reduce_mod_p(k);
            uint32_t carry = 0;
            for (uint32_t w = 0; w < m.nwin; w++) {
                uint32_t d = window_bits256(k, w, c) + carry, neg = 0;
                carry = d > half;
                if (carry) { d = full - d; neg = 1u; }
                const uint32_t wd = d ? (key_encode(d - 1u) | (neg << 31)) : ZK_KEY_NONE;
                reinterpret_cast<uint32_t*>(digits)[segs[m.seg0 + w].ent0 + i] = wd;
                atomicAdd(&cnt[w * 128 + ((wd & 0xffu) >> 1)], 1u << ((wd & 1u) * 16));
            }
            carry_left |= carry != 0;
            continue;
        }
        GlvSplit sp = glv_split(k);
        uint32_t carry1 = 0, carry2 = 0;
        uint32_t v = 0, sg = m.seg0;  // window w: local window v of segment sg
        for (uint32_t w = 0; w < m.nwin; w++) {
            uint32_t d1 = window_bits128(sp.k1, w, c) + carry1;
            uint32_t d2 = window_bits128(sp.k2, w, c) + carry2;
            uint32_t neg1 = sp.neg1 ? 1u : 0u, neg2 = sp.neg2 ? 1u : 0u;
            carry1 = d1 > half;
            carry2 = d2 > half;
            if (carry1) { d1 = full - d1; neg1 ^= 1u; }
            if (carry2) { d2 = full - d2; neg2 ^= 1u; }
            const uint32_t vt = v << 16;
            uint32_t w1 = d1 ? (key_encode(d1 - 1u) | vt | (neg1 << 31)) : ZK_KEY_NONE;
            uint32_t w2 = d2 ? (key_encode(d2 - 1u) | vt | (neg2 << 31)) : ZK_KEY_NONE;
            const uint64_t e = segs[sg].ent0 + v * vstride + 2 * i;
            digits[e >> 1] = make_uint2(w1, w2);
            atomicAdd(&cnt[w * 128 + ((w1 & 0xffu) >> 1)], 1u << ((w1 & 1u) * 16));
            atomicAdd(&cnt[w * 128 + ((w2 & 0xffu) >> 1)], 1u << ((w2 & 1u) * 16));
            if (++v == m.wps) {
                v = 0;
                sg++;
            }
        }
        carry_left |= (carry1 | carry2) != 0;
    }
    if (carry_left) atomicOr(err, 1u);
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < m.nwin * 256u; x += kDecThreads) {
        const uint32_t w = x >> 8, d = x & 255u;
        const SegDesc sg = segs[m.seg0 + w / m.wps];
        hist[(uint64_t)sg.tile0 * 256 + (uint64_t)d * sg.ntiles + (w % m.wps) * ntw + tin] =
            (cnt[w * 128 + (d >> 1)] >> ((d & 1u) * 16)) & 0xffffu;
    }
}

// ------------------------------------------------------ radix sort passes
// Two LSD passes of 8 bits over the 16-bit bucket key, independently per
// segment.  A tile is kSortTile consecutive entries of one segment; wave w of
// the block owns tile positions [w * 64 * kSortItems, (w + 1) * 64 *
// kSortItems) and walks them 64 at a time, so the order kept is (wave, round,
// lane).  Per-wave digit counters live in LDS and are only touched by their
// own wave; ranks come from LDS atomics (the first pass need not be stable,
// the second keeps the first pass's order where it matters, see
// k_sort_scatter).  Tile histograms are stored bin-major per segment,
// hist[tile0*256 + bin*ntiles + t], so one exclusive scan per segment yields
// every (bin, tile) destination.

__device__ __forceinline__ uint32_t sort_key(const void* in, int pass, uint64_t pos) {
    if (pass == 0) return reinterpret_cast<const uint32_t*>(in)[pos] & 0xffffu;
    return reinterpret_cast<const uint16_t*>(in)[pos];
}

// Entries of a tile: pass 0 sorts all 2n of the segment; pass 1 only the
// seg_valid[s] entries with a bucket (pass 0 put them first).
__device__ __forceinline__ uint32_t tile_len(const SegDesc& sg, uint32_t start, int pass,
                                             const uint32_t* seg_valid, uint32_t s) {
    const uint32_t n = pass ? seg_valid[s] : sg.len;
    return start < n ? min((uint32_t)kSortTile, n - start) : 0u;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const SegDesc* __restrict__ segs,
                                                            const uint32_t* __restrict__ tile_seg,
                                                            uint32_t tile_lo,
                                                            const void* __restrict__ in, int pass,
                                                            const uint32_t* __restrict__ seg_valid,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kSortWaves][256];
    const uint32_t t = tile_lo + blockIdx.x;
    const uint32_t s = tile_seg[t];
    const SegDesc sg = segs[s];
    const uint32_t tin = t - sg.tile0;
    const uint32_t start = tin * kSortTile;
    const uint32_t len = tile_len(sg, start, pass, seg_valid, s);
    const uint32_t shift = pass ? 8u : 0u;
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    for (uint32_t i = threadIdx.x; i < kSortWaves * 256u; i += kSortThreads) (&cnt[0][0])[i] = 0;
    __syncthreads();
    // counts only (no ranks): one LDS atomic per entry into the wave's own
    // counters is far cheaper than matching equal digits with ballots
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t pos = w * (64u * kSortItems) + r * 64u + lane;
        if (pos < len) atomicAdd(&cnt[w][(sort_key(in, pass, sg.ent0 + start + pos) >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        const uint32_t d = threadIdx.x;
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < kSortWaves; i++) s += cnt[i][d];
        hist[(uint64_t)sg.tile0 * 256 + (uint64_t)d * sg.ntiles + tin] = s;
    }
}

// Exclusive scan of each segment's 256*ntiles counts into global destinations
// (+ ent0): reduce-then-scan over chunks of kScanChunk counts -- up (chunk
// sums), spine (per segment, over chunk sums), down (rescan + offsets).  Pass 1
// also records the number of entries with a real bucket: the exclusive
// prefix of bin 0xFF, which only the "no bucket" key 0xFFFF reaches.

__device__ __forceinline__ void load16(const uint32_t* p, uint32_t v[16]) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint4 x = q[i];
        v[4 * i] = x.x;
        v[4 * i + 1] = x.y;
        v[4 * i + 2] = x.z;
        v[4 * i + 3] = x.w;
    }
}

__global__ __launch_bounds__(256) void k_scan_up(const SegDesc* __restrict__ segs,
                                                 const uint32_t* __restrict__ chunk_seg, uint32_t chunk_lo,
                                                 const uint32_t* __restrict__ hist,
                                                 uint32_t* __restrict__ part) {
    __shared__ uint32_t lds4[4];
    const uint32_t c = chunk_lo + blockIdx.x;
    const SegDesc sg = segs[chunk_seg[c]];
    const uint32_t M = 256u * sg.ntiles;
    const uint32_t i0 = (c - sg.chunk0) * kScanChunk + 16u * threadIdx.x;
    uint32_t s = 0;
    if (i0 < M) {
        uint32_t v[16];
        load16(hist + (uint64_t)sg.tile0 * 256 + i0, v);
#pragma unroll
        for (int i = 0; i < 16; i++) s += v[i];
    }
    uint32_t tot;
    (void)block_exclusive_scan(s, lds4, &tot);
    if (threadIdx.x == 0) part[c] = tot;
}

__global__ __launch_bounds__(256) void k_scan_spine(const SegDesc* __restrict__ segs, uint32_t seg_lo,
                                                    uint32_t* __restrict__ part) {
    __shared__ uint32_t lds4[4];
    const SegDesc sg = segs[seg_lo + blockIdx.x];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < sg.nchunks; b += 256) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < sg.nchunks ? part[sg.chunk0 + i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, lds4, &tot);
        if (i < sg.nchunks) part[sg.chunk0 + i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(256) void k_scan_down(const SegDesc* __restrict__ segs,
                                                   const uint32_t* __restrict__ chunk_seg, uint32_t chunk_lo,
                                                   uint32_t* __restrict__ hist,
                                                   const uint32_t* __restrict__ part, int pass,
                                                   uint32_t* __restrict__ seg_valid) {
    __shared__ uint32_t lds4[4];
    const uint32_t c = chunk_lo + blockIdx.x;
    const uint32_t s = chunk_seg[c];
    const SegDesc sg = segs[s];
    const uint32_t M = 256u * sg.ntiles;
    const uint32_t i0 = (c - sg.chunk0) * kScanChunk + 16u * threadIdx.x;
    uint32_t v[16];
    uint32_t sum = 0;
    uint32_t* h = hist + (uint64_t)sg.tile0 * 256 + i0;
    if (i0 < M) {
        load16(h, v);
#pragma unroll
        for (int i = 0; i < 16; i++) sum += v[i];
    }
    uint32_t tot;
    uint32_t acc = part[c] + block_exclusive_scan(sum, lds4, &tot);
    if (i0 < M) {
        const uint32_t none_bin = 0xffu * sg.ntiles;
        uint32_t o[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if (pass == 0 && i0 + i == none_bin) seg_valid[s] = acc;
            o[i] = (uint32_t)sg.ent0 + acc;
            acc += v[i];
        }
        uint4* q = reinterpret_cast<uint4*>(h);
#pragma unroll
        for (int i = 0; i < 4; i++) q[i] = make_uint4(o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]);
    }
}

// Scatter: rank every entry of the tile (wave-private counters; stable ranks
// from ballot matching in the second pass), sort the tile in LDS, then write
// each digit's run to its global destination with consecutive threads
// writing consecutive entries.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const SegDesc* __restrict__ segs, const uint32_t* __restrict__ tile_seg, uint32_t tile_lo,
    const void* __restrict__ in_keys, const uint32_t* __restrict__ in_vals, int pass,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ seg_valid,
    uint16_t* __restrict__ out_keys, uint32_t* __restrict__ out_vals) {
    __shared__ uint32_t cnt[kSortWaves][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t goff[256];
    __shared__ uint32_t ldsw[kSortWaves];
    __shared__ uint16_t lkey[kSortTile];
    __shared__ uint32_t lval[kSortTile];
    const uint32_t t = tile_lo + blockIdx.x;
    const uint32_t s = tile_seg[t];
    const SegDesc sg = segs[s];
    const uint32_t tin = t - sg.tile0;
    const uint32_t start = tin * kSortTile;
    const uint32_t len = tile_len(sg, start, pass, seg_valid, s);
    const uint32_t shift = pass ? 8u : 0u;
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    for (uint32_t i = threadIdx.x; i < kSortWaves * 256u; i += kSortThreads) (&cnt[0][0])[i] = 0;
    __syncthreads();
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t pos = w * (64u * kSortItems) + r * 64u + lane;
        const bool act = pos < len;
        key[r] = ZK_KEY_NONE;
        val[r] = 0;
        if (act) {
            const uint64_t gp = sg.ent0 + start + pos;
            if (pass == 0) {
                const uint32_t wd = reinterpret_cast<const uint32_t*>(in_keys)[gp];
                key[r] = wd & 0xffffu;
                // the table entry: the position, shifted to its window's copy
                val[r] = (start + pos + ((wd >> 16) & 0xffu) * sg.wadj + sg.wbase) | (wd & 0x80000000u);
            } else {
                key[r] = reinterpret_cast<const uint16_t*>(in_keys)[gp];
                val[r] = in_vals[gp];
            }
        }
        const uint32_t digit = (key[r] >> shift) & 0xffu;
        if (pass == 0) {
            // the first pass need not be stable (the second, stable pass fixes
            // the final order): ranks from LDS atomics, no digit matching
            rank[r] = act ? atomicAdd(&cnt[w][digit], 1u) : 0u;
        } else {
            // The second pass must keep the first pass's order between entries
            // of different low bytes; entries with the same low byte and the
            // same high byte share a bucket, so their order is free.  The 64
            // positions of one instruction are consecutive in that order, so
            // their low bytes ascend with the lane: the lanes of one low byte
            // at a time take LDS-atomic ranks, lowest low byte first (about
            // two rounds per instruction for uniform digits).
            const uint32_t lowb = key[r] & 0xffu;
            bool todo = act;
            uint32_t rk = 0;
            for (uint64_t m = __ballot(todo); m != 0; m = __ballot(todo)) {
                const uint32_t lb = __builtin_amdgcn_readlane(lowb, __ffsll((long long)m) - 1);
                if (todo && lowb == lb) {
                    rk = atomicAdd(&cnt[w][digit], 1u);
                    todo = false;
                }
            }
            rank[r] = rk;
        }
    }
    __syncthreads();
    {
        // digit totals by threads 0..255; per-wave starts within each digit
        const uint32_t d = threadIdx.x;
        uint32_t s = 0;
        if (d < 256) {
#pragma unroll
            for (int i = 0; i < kSortWaves; i++) s += cnt[i][d];
        }
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<kSortWaves>(s, ldsw, &tot);
        if (d < 256) {
            dstart[d] = ex;
            uint32_t run = ex;
#pragma unroll
            for (int i = 0; i < kSortWaves; i++) {
                const uint32_t c = cnt[i][d];
                cnt[i][d] = run;
                run += c;
            }
            goff[d] = hist[(uint64_t)sg.tile0 * 256 + (uint64_t)d * sg.ntiles + tin];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t pos = w * (64u * kSortItems) + r * 64u + lane;
        if (pos < len) {
            const uint32_t j = cnt[w][(key[r] >> shift) & 0xffu] + rank[r];
            lkey[j] = (uint16_t)key[r];
            lval[j] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < len; j += kSortThreads) {
        const uint32_t k = lkey[j];
        const uint32_t d = (k >> shift) & 0xffu;
        const uint32_t g = goff[d] + (j - dstart[d]);
        if (k != ZK_KEY_NONE) {  // entries without a bucket are not kept
            out_keys[g] = (uint16_t)k;
            out_vals[g] = lval[j];
        }
    }
}

// ------------------------------------------------------------ utilities

__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ p, uint64_t n16) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

// --------------------------------------------------- bucket accumulation
// Lane L of segment s (l = L - seg_lane0[s]; T and seg_lane0 as k_lane_plan
// set them for the part) owns sorted entries [ent0 + l*T, ent0 + (l+1)*T)
// clipped to the segment's valid entries.  It walks them once with mixed additions;
// a bucket run that starts and ends inside the range is final and is written
// to the bucket array, the (at most two) runs that cross the range ends go to
// merge records 2L (first run) and 2L+1 (last run).

template <class C>
__device__ __forceinline__ void put_record(uint32_t* keys, uint4* pts, uint64_t slot,
                                           uint32_t key, bool valid, const typename C::Pt& v) {
    keys[slot] = key | (valid ? 0u : ZK_REC_INVALID);
    if (valid) C::store(pts, slot, v);
}

// Lanes of one part, [seg_lo, seg_lo + nseg), from the sort's counts of
// entries that have a bucket (seg_valid): the planner sizes T for all of a
// segment's entries, but the sort drops the zero digits -- 70% of them for
// witness-like scalars -- and the lanes past a segment's valid entries have
// nothing to do.  Here T = max(4, min(32, half a bucket's mean valid load),
// valid / (lanes - nseg)), so the part's planned lanes (one resident round)
// are all busy again, and seg_lane0[s] = lane_lo + sum over the earlier
// segments of ceil(valid / T) (at most lanes in all: each segment rounds up
// by less than one lane).  One workgroup; T goes to *Tout for k_accumulate.
__global__ __launch_bounds__(256) void k_lane_plan(const SegDesc* __restrict__ segs,
                                                   const uint32_t* __restrict__ seg_valid, uint32_t seg_lo,
                                                   uint32_t nseg, uint64_t lane_lo, uint64_t lanes,
                                                   uint64_t* __restrict__ seg_lane0, uint32_t* __restrict__ Tout) {
    __shared__ uint32_t lds4[4];
    uint32_t v = 0, b = 0;
    for (uint32_t i = threadIdx.x; i < nseg; i += 256) {
        v += seg_valid[seg_lo + i];
        b += segs[seg_lo + i].nbk;
    }
    uint32_t V, B;
    (void)block_exclusive_scan(v, lds4, &V);
    (void)block_exclusive_scan(b, lds4, &B);
    const uint64_t cap = lanes > nseg ? lanes - nseg : 1;
    uint32_t T = max(4u, min(32u, V / (2u * max(B, 1u))));
    T = max(T, (uint32_t)((V + cap - 1) / cap));
    if (threadIdx.x == 0) *Tout = T;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < nseg; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t c = i < nseg ? (seg_valid[seg_lo + i] + T - 1) / T : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(c, lds4, &tot);
        if (i < nseg) seg_lane0[seg_lo + i] = lane_lo + carry + ex;
        carry += tot;
    }
}

template <class C>
__global__ __launch_bounds__(256) void k_accumulate(
    const SegDesc* __restrict__ segs, const uint64_t* __restrict__ seg_lane0, uint32_t seg_lo, uint32_t nseg,
    const MsmDesc* __restrict__ msms, const uint32_t* __restrict__ seg_valid,
    const uint16_t* __restrict__ keys, const uint32_t* __restrict__ vals,
    const uint4* __restrict__ table, const uint32_t* __restrict__ Tp, uint64_t lane_lo, uint64_t lane_hi,
    uint4* __restrict__ buckets, uint32_t* __restrict__ rec_keys, uint4* __restrict__ rec_pts) {
    const uint64_t L = lane_lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= lane_hi) return;
    const uint32_t T = *Tp;  // k_lane_plan's entries per lane for this part
    const uint32_t s = seg_lo + upper_index(seg_lane0 + seg_lo, nseg, L);
    const SegDesc sg = segs[s];
    const uint64_t l = L - seg_lane0[s];
    const uint64_t segend = sg.ent0 + seg_valid[s];
    const uint64_t b = sg.ent0 + l * T;
    const uint64_t e = min(b + T, segend);
    if (b >= e) {
        rec_keys[2 * L] = ZK_REC_INVALID | ZK_REC_KEYMASK;
        rec_keys[2 * L + 1] = ZK_REC_INVALID | ZK_REC_KEYMASK;
        return;
    }
    const uint4* tb = table + msms[sg.msm].table0 * C::kSlots * (kTableEntryBytes / 16);
    uint32_t cur = keys[b];
    const bool head_open = (b > sg.ent0) && keys[b - 1] == cur;
    bool first = true;
    typename C::Pt acc = C::run_start();
    bool acc_id = true;
    // software pipeline: before entry p's mixed addition the base of entry
    // p+1 is gathered (its value was loaded one iteration earlier, so the
    // gather's address waits on nothing issued in this iteration) and the key
    // and value of entry p+2 are loaded; indices clamp to the lane's last entry
    uint32_t k = cur;
    const uint32_t v0 = vals[b];
    typename C::Af q = C::load_tab(tb, v0 & 0x7fffffffu, v0 >> 31);
    const uint64_t p1 = min(b + 1, e - 1);
    uint32_t kn = keys[p1], vn = vals[p1];
    for (uint64_t p = b; p < e; p++) {
        const uint64_t pnn = min(p + 2, e - 1);
        const uint32_t knn = keys[pnn], vnn = vals[pnn];
        const typename C::Af qn = C::load_tab(tb, vn & 0x7fffffffu, vn >> 31);
        if (k != cur) {
            // run `cur` ended at p-1 and is not the lane's last run
            const uint32_t gk = (uint32_t)(sg.bkt0 + key_decode(cur));
            if (first && head_open) {
                put_record<C>(rec_keys, rec_pts, 2 * L, gk, true, acc);
            } else {
                C::store(buckets, gk, acc);
                if (first) put_record<C>(rec_keys, rec_pts, 2 * L, gk, false, acc);
            }
            first = false;
            acc = C::run_start();
            acc_id = true;
            cur = k;
        }
        // y is already negated for a negative digit; identity bases have no
        // digits (table bitmap), so q is never the identity
        C::madd_run(acc, acc_id, q);
        k = kn;
        kn = knn;
        vn = vnn;
        q = qn;
    }
    const bool tail_open = (e < segend) && keys[e] == cur;
    const uint32_t gk = (uint32_t)(sg.bkt0 + key_decode(cur));
    if (first) {
        const bool open = head_open || tail_open;
        if (!open) C::store(buckets, gk, acc);
        put_record<C>(rec_keys, rec_pts, 2 * L, gk, open, acc);
        put_record<C>(rec_keys, rec_pts, 2 * L + 1, gk, false, acc);
    } else {
        if (!tail_open) C::store(buckets, gk, acc);
        put_record<C>(rec_keys, rec_pts, 2 * L + 1, gk, tail_open, acc);
    }
}

// ------------------------------------------------------ partial merging
// Records with equal key are contiguous (a bucket's partials come from
// consecutive lanes; padding of empty lanes carries key 0x7fffffff).
// k_merge_runs: the lane holding the first record of a run with at most
// `limit` payload records (and at most 3*limit records in all) sums it and
// writes dest[key]; a run of at most 64 records is appended to mid_list, a
// longer one to long_list.  k_merge_mid: eight lanes per listed run (eight
// runs per wave), each summing every eighth record, then a 3-level shuffle
// tree.  k_merge_long: one 256-lane workgroup per listed run, a shuffle tree
// per wave, then one lane adds the four wave partials: depth len/256 + 9.

template <class C>
__global__ __launch_bounds__(256) void k_merge_runs(const uint32_t* __restrict__ keys,
                                                    const uint4* __restrict__ pts, uint64_t nrec,
                                                    uint32_t limit, uint4* __restrict__ dest,
                                                    uint32_t* __restrict__ long_count,
                                                    uint32_t* __restrict__ long_list,
                                                    uint32_t* __restrict__ mid_count,
                                                    uint32_t* __restrict__ mid_list,
                                                    uint32_t list_cap) {
    // lane i looks at records 2i and 2i+1 (one accumulation lane's head and
    // tail slots); at most one of them starts a run with a payload
    const uint64_t i2 = 2 * ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i2 >= nrec) return;
    for (uint64_t i = i2; i < i2 + 2 && i < nrec; i++) {
        const uint32_t k = keys[i] & ZK_REC_KEYMASK;
        if (k == ZK_REC_KEYMASK) continue;  // padding
        if (i > 0 && (keys[i - 1] & ZK_REC_KEYMASK) == k) continue;
        uint64_t j = i;
        uint32_t valid = 0;
        while (j < nrec && valid <= limit && j - i <= 3 * limit) {
            const uint32_t kj = keys[j];
            if ((kj & ZK_REC_KEYMASK) != k) break;
            valid += (kj & ZK_REC_INVALID) ? 0u : 1u;
            j++;
        }
        if (valid > limit || (j - i > 3 * limit && j < nrec && (keys[j] & ZK_REC_KEYMASK) == k)) {
            // records are sorted by key: the run is longer than 64 records
            // iff record i + 64 still carries its key
            const bool is_long = i + 64 < nrec && (keys[i + 64] & ZK_REC_KEYMASK) == k;
            const uint32_t slot = atomicAdd(is_long ? long_count : mid_count, 1u);
            if (slot < list_cap) (is_long ? long_list : mid_list)[slot] = (uint32_t)i;
            continue;
        }
        bool has = false;
        typename C::Pt acc = C::identity();
        for (uint64_t p = i; p < j; p++) {
            if (!(keys[p] & ZK_REC_INVALID)) {
                typename C::Pt q = C::load(pts, p);
                if (has) C::add(acc, q);
                else acc = q;
                has = true;
            }
        }
        if (has) C::store(dest, k, acc);
    }
}

template <class C>
__global__ __launch_bounds__(256) void k_merge_mid(const uint32_t* __restrict__ keys,
                                                   const uint4* __restrict__ pts, uint64_t nrec,
                                                   const uint32_t* __restrict__ mid_count,
                                                   const uint32_t* __restrict__ mid_list,
                                                   uint32_t list_cap, uint4* __restrict__ dest) {
    const uint32_t n = min(*mid_count, list_cap);
    const uint32_t lane = __lane_id(), sub = lane & 7u;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    for (uint64_t base = wave * 8; base < n; base += nwaves * 8) {  // uniform per wave
        const uint64_t run = base + (lane >> 3);
        uint64_t start = 0;
        uint32_t k = 0, cnt = 0;
        bool has = false;
        if (run < n) {
            start = mid_list[run];
            k = keys[start] & ZK_REC_KEYMASK;
            for (uint64_t p = start + sub; p < start + 64 && p < nrec; p += 8) {
                const uint32_t kr = keys[p];
                if ((kr & ZK_REC_KEYMASK) != k) break;
                cnt++;
                has |= !(kr & ZK_REC_INVALID);
            }
        }
        const typename C::Pt acc = sum_then_tree<C>(wave_max(cnt), 8, [&](uint32_t e) {
            const uint64_t p = start + sub + 8ull * e;
            return load_or_id<C>(pts, p, e < cnt && !(keys[p] & ZK_REC_INVALID));
        });
        // the group's first lane stores if any lane of the group had a payload
        const uint64_t hb = __ballot(has);
        if (run < n && sub == 0 && ((hb >> (lane & ~7u)) & 0xffull)) C::store(dest, k, acc);
    }
}

template <class C>
__global__ __launch_bounds__(256) void k_merge_long(const uint32_t* __restrict__ keys,
                                                    const uint4* __restrict__ pts, uint64_t nrec,
                                                    const uint32_t* __restrict__ long_count,
                                                    const uint32_t* __restrict__ long_list,
                                                    uint32_t long_cap, uint4* __restrict__ dest) {
    __shared__ uint4 part[4][C::kPtBytes / 16];
    __shared__ uint32_t part_has[4];
    const uint32_t n = min(*long_count, long_cap);
    const uint32_t t = threadIdx.x, wv = t >> 6, lane = __lane_id();
    for (uint32_t run = blockIdx.x; run < n; run += gridDim.x) {  // uniform per workgroup
        const uint64_t start = long_list[run];
        const uint32_t k = keys[start] & ZK_REC_KEYMASK;
        uint32_t cnt = 0;
        bool has = false;
        for (uint64_t p = start + t; p < nrec; p += 256) {
            const uint32_t kr = keys[p];
            if ((kr & ZK_REC_KEYMASK) != k) break;
            cnt++;
            has |= !(kr & ZK_REC_INVALID);
        }
        typename C::Pt acc = sum_then_tree<C>(wave_max(cnt), 64, [&](uint32_t e) {
            const uint64_t p = start + t + 256ull * e;
            return load_or_id<C>(pts, p, e < cnt && !(keys[p] & ZK_REC_INVALID));
        });
        const bool wave_has = __ballot(has) != 0;
        if (lane == 0) {
            C::store(&part[wv][0], 0, acc);
            part_has[wv] = wave_has ? 1u : 0u;
        }
        __syncthreads();
        if (t == 0) {
#pragma unroll 1
            for (uint32_t w = 1; w < 4; w++) {
                const typename C::Pt o = C::load(&part[w][0], 0);
                C::add(acc, o);
            }
            if (part_has[0] | part_has[1] | part_has[2] | part_has[3]) C::store(dest, k, acc);
        }
        __syncthreads();  // LDS partials are reused by the next run
    }
}

// ------------------------------------------------------ bucket reduction
// Window sum S = sum_k (k+1) B_k over bucket keys k < 2^(c-1).  Write
// k = hi * 2^g + lo (g = (c-1)/2 low bits, h = c-1-g high bits).  Then
//   S = U + sum_{b<g} 2^b Pl_b + sum_{b<h} 2^(g+b) Ph_b,
//   U = sum_lo C_lo,  Pl_b = sum_{lo: bit b} C_lo,  Ph_b = sum_{hi: bit b} R_hi,
// with row sums R_hi = sum_lo B_{hi,lo} and column sums C_lo = sum_hi B_{hi,lo}
// (two additions per bucket in all).  k_br_rowcol: a row of 2^g buckets is
// summed by P_r = max(1, 2^g/per) lanes (`per` buckets each, then a
// log2(P_r) shuffle tree), a column of 2^h buckets by P_c = max(1, 2^h/per)
// lanes; rows and columns of a segment start on wave boundaries so P is
// wave-uniform.  per = 8..32 is chosen by the planner from the call's size.
// k_br_bits: one wave per output point, bits[s*16 + j]: j = 0 -> U,
// 1..g -> Pl_{j-1}, g+1..c-1 -> Ph_{j-1-g}.  The host applies the powers of 2.

__device__ __forceinline__ void seg_shape(uint32_t nbk, uint32_t& g, uint32_t& h) {
    const uint32_t cb = 31 - __clz(nbk);  // c - 1
    g = cb >> 1;
    h = cb - g;
}

__device__ __forceinline__ uint32_t lanes_per(uint32_t len, uint32_t per) { return len >= per ? len / per : 1; }

template <class C>
__global__ __launch_bounds__(256) void k_br_rowcol(const SegDesc* __restrict__ segs,
                                                   const uint64_t* __restrict__ seg_wave0,
                                                   const uint64_t* __restrict__ seg_rc0,
                                                   uint32_t nseg, uint64_t wave_lo, uint64_t wave_hi,
                                                   uint32_t per, const uint4* __restrict__ buckets,
                                                   uint4* __restrict__ rc) {
    const uint64_t wave = wave_lo + (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= wave_hi) return;
    const uint32_t lane = __lane_id();
    const uint32_t s = upper_index(seg_wave0, nseg, wave);
    const SegDesc sg = segs[s];
    uint32_t g, h;
    seg_shape(sg.nbk, g, h);
    const uint32_t nrow = 1u << h, ncol = 1u << g;
    const uint32_t Pr = lanes_per(ncol, per), Pc = lanes_per(nrow, per);
    const uint32_t row_waves = (nrow * Pr + 63) / 64;
    uint32_t w = (uint32_t)(wave - seg_wave0[s]);
    const bool rows = w < row_waves;
    if (!rows) w -= row_waves;
    const uint32_t P = rows ? Pr : Pc;
    const uint32_t li = w * 64 + lane;
    const uint32_t item = li / P, part = li % P;
    const uint32_t nitem = rows ? nrow : ncol;
    const uint32_t len = rows ? ncol : nrow;  // buckets per item
    const uint32_t cnt = len / P;  // buckets this lane sums
    const typename C::Pt acc = sum_then_tree<C>(cnt, P, [&](uint32_t e) {
        const uint32_t x = part * cnt + e;
        const uint64_t kk = rows ? ((uint64_t)item << g) + x : ((uint64_t)x << g) + item;
        return load_or_id<C>(buckets, sg.bkt0 + kk, item < nitem);
    });
    if (item < nitem && part == 0) C::store(rc, seg_rc0[s] + (rows ? item : nrow + item), acc);
}

// index of the i-th value with bit b set: insert a 1 at bit position b
__device__ __forceinline__ uint32_t with_bit(uint32_t i, uint32_t b) {
    return ((i >> b) << (b + 1)) | (1u << b) | (i & ((1u << b) - 1u));
}

template <class C>
__global__ __launch_bounds__(256) void k_br_bits(const SegDesc* __restrict__ segs,
                                                 const uint64_t* __restrict__ seg_rc0,
                                                 uint32_t seg_lo, uint32_t seg_hi, const uint4* __restrict__ rc,
                                                 uint4* __restrict__ bits, int host_form) {
    const uint64_t item = (uint64_t)seg_lo * 16 + (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t s = (uint32_t)(item >> 4), j = (uint32_t)(item & 15u);
    if (s >= seg_hi) return;
    const uint32_t lane = __lane_id();
    const SegDesc sg = segs[s];
    uint32_t g, h;
    seg_shape(sg.nbk, g, h);
    if (j > g + h) return;
    const uint64_t R0 = seg_rc0[s], C0 = R0 + (1u << h);
    // element list: U -> all columns; Pl_b -> columns with bit b; Ph_b -> rows with bit b
    uint64_t base;
    uint32_t count, b = 0;
    bool masked;
    if (j == 0) {
        base = C0; count = 1u << g; masked = false;
    } else if (j <= g) {
        base = C0; b = j - 1; count = 1u << (g - 1); masked = true;
    } else {
        base = R0; b = j - 1 - g; count = 1u << (h - 1); masked = true;
    }
    // count is a power of 2: lanes take elements lane, lane+64, ...
    const uint32_t P = count >= 64 ? 64u : count;
    const typename C::Pt acc = sum_then_tree<C>((count + 63) / 64, P, [&](uint32_t e) {
        const uint32_t i = lane + 64 * e;
        return load_or_id<C>(rc, base + (masked ? with_bit(i, b) : i), i < count);
    });
    if (lane == 0) {
        if (host_form) C::store_host(bits, item, acc);  // for the host tail
        else C::store(bits, item, acc);                 // for k_tail_windows
    }
}

// ------------------------------------------------ window combine on the device
// The device form of the call's tail (the host form is Engine::run's
// host_tail): result_j = sum_w 2^(cstep w) S_w over the MSM's segments w,
// S_w = U_w + sum_{t < c-1} 2^t Q_{w,t}, the planes k_br_bits stored
// (bits[s*16] = U, bits[s*16 + 1 + t] = Q_t), then the affine normalisation;
// cstep = c for plain windows, 0 for window-shifted segments (a plain sum).
// The reference does its window combine on the host
// (Sources/zkMetal/MSM/MSMEngine.swift:899-904); on the device it lets a
// captured graph end with final results.  The doubling chain is inherently
// serial (2^127 S_top needs 127 doublings), so sixteen lanes per segment
// compute S_w, then one lane per MSM runs the Horner over segments and the
// inversion.

// k_tail_windows: sixteen lanes per segment (four segments per wave): lane
// t < c - 1 doubles plane Q_t t times (a Jacobian chain), lane 15 holds U,
// then a 4-level shuffle tree sums the sixteen.  The longest chain is c - 2
// doublings and four additions, against c - 2 doublings and c - 1 additions
// one after another in a Horner over the planes.
template <class C>
__global__ __launch_bounds__(64) void k_tail_windows(const SegDesc* __restrict__ segs,
                                                     const MsmDesc* __restrict__ msms, uint32_t nseg,
                                                     const uint4* __restrict__ bits,
                                                     uint4* __restrict__ wins) {
    const uint32_t t = threadIdx.x & 15u;
    const uint32_t sidx = (blockIdx.x * 64 + threadIdx.x) >> 4;
    const bool valid = sidx < nseg;
    const uint32_t s = valid ? sidx : 0u;
    const uint32_t c = msms[segs[s].msm].c;
    const uint64_t b0 = (uint64_t)s * 16;
    typename C::Pt v = C::identity();
    if (valid && t + 1 < c) v = C::dbl_k(C::load(bits, b0 + 1 + t), t);  // 2^t Q_t
    if (valid && t == 15) v = C::load(bits, b0);                         // U (c - 1 <= 15 planes)
#pragma unroll
    for (uint32_t d = 8; d >= 1; d >>= 1) {
        const typename C::Pt o = C::shfl_down(v, d);
        if (t < d) C::add(v, o);
    }
    if (valid && t == 0) C::store(wins, s, v);
}

template <class C>
__global__ __launch_bounds__(64) void k_tail_combine(const MsmDesc* __restrict__ msms, uint32_t nmsm,
                                                     const uint4* __restrict__ wins,
                                                     uint4* __restrict__ out) {
    const uint32_t j = blockIdx.x * 64 + threadIdx.x;
    if (j >= nmsm) return;
    const MsmDesc m = msms[j];
    typename C::Pt acc = C::load(wins, m.seg0 + m.nseg - 1);
#pragma unroll 1
    for (int w = (int)m.nseg - 2; w >= 0; w--) {
        // 2^c acc, in Jacobian coordinates (window-shifted segments: cstep = 0,
        // their sums already carry the window weights)
        if (m.cstep) acc = C::dbl_k(acc, m.cstep);
        C::add(acc, C::load(wins, m.seg0 + w));
    }
    C::store_result(out, j, acc);
}

// ------------------------------------------------ synthetic base table
// table[t0 + 2i] = k_i * G, table[t0 + 2i + 1] = phi(k_i * G), from a host
// table gtab[w*256 + b] = b * 2^(8w) * G (affine, entry 0 = (0,0)).

__global__ __launch_bounds__(256) void k_gen_bases(const uint4* __restrict__ ks,
                                                   const uint4* __restrict__ gtab, uint64_t n,
                                                   uint4* __restrict__ table,
                                                   uint64_t* __restrict__ idmask) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 s0 = ks[2 * i], s1 = ks[2 * i + 1];
    uint32_t k[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    Xyzz acc = xyzz_identity();
    for (int w = 0; w < 32; w++) {
        uint32_t byte = (k[w >> 2] >> ((w & 3) * 8)) & 0xffu;
        if (byte) {
            Aff q = load_aff(gtab, (uint64_t)w * 256 + byte);
            xyzz_madd(acc, q.x, q.y, false);
        }
    }
    Aff a;
    if (xyzz_is_id(acc)) {
        a.x = fp_zero();
        a.y = fp_zero();
    } else {
        Fp t = fp_inv(fp_mul(acc.ZZ, acc.ZZZ));
        a.x = fp_mul(acc.X, fp_mul(t, acc.ZZZ));
        a.y = fp_mul(acc.Y, fp_mul(t, acc.ZZ));
    }
    store_table_pair(table, idmask, i, a);
}

// Affine points out[i] = k_i * G (64 B: x, y in the 2^256 Montgomery form,
// (0,0) for k_i = 0) on the 29-bit arithmetic of field F, from a host table
// gtab[w*256 + b] = b * 2^(8w) * G in the same form (entry 0 = (0,0)):
// synthetic Grumpkin bases for benchmarks and tests.
template <class F>
__global__ __launch_bounds__(256) void k_gen_points(const uint4* __restrict__ ks,
                                                    const uint4* __restrict__ gtab, uint64_t n,
                                                    uint4* __restrict__ out) {
    using A = Arith<F>;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 s0 = ks[2 * i], s1 = ks[2 * i + 1];
    const uint32_t k[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    QXyzz acc = qxyzz_identity();
    for (int w = 0; w < 32; w++) {
        const uint32_t byte = (k[w >> 2] >> ((w & 3) * 8)) & 0xffu;
        if (byte) {
            const Aff q = load_aff(gtab, (uint64_t)w * 256 + byte);
            A::qxyzz_madd(acc, A::fq_from_fp_canon(q.x), A::fq_from_fp_canon(q.y), false);
        }
    }
    Aff a;
    if (A::qxyzz_is_id(acc)) {
        a.x = fp_zero();
        a.y = fp_zero();
    } else {
        // 1/ZZ = ZZZ t and 1/ZZZ = ZZ t with t = 1/(ZZ ZZZ)
        const Fq t = A::fq_inv(A::fq_mul(acc.ZZ, acc.ZZZ));
        a.x = A::fq_to_fp(A::fq_mul(acc.X, A::fq_mul(t, acc.ZZZ)));
        a.y = A::fq_to_fp(A::fq_mul(acc.Y, A::fq_mul(t, acc.ZZ)));
    }
    store_aff(out, i, a);
}

// BN254 G2 points out[i] = k_i * G2 (128 B: x.c0, x.c1, y.c0, y.c1 in the
// 2^256 form; all zero for k_i = 0) from a host table gtab[w*256 + b] =
// b * 2^(8w) * G2 in the same form (entry 0 all zero): synthetic G2 bases.
__global__ __launch_bounds__(256) void k_gen_points_g2(const uint4* __restrict__ ks,
                                                       const uint4* __restrict__ gtab, uint64_t n,
                                                       uint4* __restrict__ out) {
    using G = G2Arith;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 s0 = ks[2 * i], s1 = ks[2 * i + 1];
    const uint32_t k[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    QXyzz2 acc = G::identity();
    for (int w = 0; w < 32; w++) {
        const uint32_t byte = (k[w >> 2] >> ((w & 3) * 8)) & 0xffu;
        if (byte) {
            const uint4* q = gtab + ((uint64_t)w * 256 + byte) * 8;
            G::madd(acc, G::from_fp2(load_fp(q), load_fp(q + 2)), G::from_fp2(load_fp(q + 4), load_fp(q + 6)), false);
        }
    }
    uint4* o = out + i * 8;
    if (G::is_id(acc)) {
        for (int j = 0; j < 8; j++) o[j] = make_uint4(0, 0, 0, 0);
        return;
    }
    // 1/ZZ = ZZZ t and 1/ZZZ = ZZ t with t = 1/(ZZ ZZZ)
    const Fq2 t = G::inv2(G::mul2(acc.ZZ, acc.ZZZ));
    const Fq2 x = G::mul2(acc.X, G::mul2(t, acc.ZZZ));
    const Fq2 y = G::mul2(acc.Y, G::mul2(t, acc.ZZ));
    store_fp(o, G::to_fp(x.c0));
    store_fp(o + 2, G::to_fp(x.c1));
    store_fp(o + 4, G::to_fp(y.c0));
    store_fp(o + 6, G::to_fp(y.c1));
}

// BN254 G2 primitives on host-form XYZZ points (256 B: c0 record, c1 record).
// op 0: P + Q, 1: 2P, 2: P + affine (Q.X, Q.Y) (all-zero = identity), 3: the
// run addition P + (Q.X, Q.Y) from a run start (acc_id) or not; field ops on
// the X coordinate: 4: X_P * X_Q, 5: X_P^2, 6: 1 / X_P (result in X, the rest
// zero).
__global__ __launch_bounds__(64) void k_test_g2(int op, const uint4* __restrict__ p,
                                                const uint4* __restrict__ q, uint4* __restrict__ out,
                                                uint64_t n) {
    using G = G2Arith;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    QXyzz2 a = load_host_xyzz2(p, i), b = load_host_xyzz2(q, i);
    const bool b_aff_id = G::is_zero2(b.X) && G::is_zero2(b.Y);
    if (op == 0) G::add(a, b);
    else if (op == 1) a = G::dbl(a);
    else if (op == 2) G::madd(a, b.X, b.Y, b_aff_id);
    else if (op == 3) {
        bool id = G::is_id(a);
        if (id) a = G::run_start();
        G::madd_run(a, id, b.X, b.Y);
        if (id) a = G::identity();
    } else {
        const Fq2 r = op == 4 ? G::mul2(a.X, b.X) : op == 5 ? G::sqr2(a.X) : G::inv2(a.X);
        a = G::identity();
        a.X = r;
    }
    store_host_xyzz2(out, i, a);
}

// ------------------------------------------------------ test entry points
// Field / curve primitives exposed for the GPU unit tests.

__global__ __launch_bounds__(64) void k_test_fp(int op, const uint4* __restrict__ a, const uint4* __restrict__ b,
                          uint4* __restrict__ out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fp x = load_fp(a + 2 * i), y = load_fp(b + 2 * i), r;
    switch (op) {
        case 0: r = fp_mul(x, y); break;
        case 1: r = fp_add(x, y); break;
        case 2: r = fp_sub(x, y); break;
        case 3: r = fp_inv(x); break;
        case 4: r = fp_neg(x); break;
        case 6: r = Ap::fq_to_fp(Ap::fq_mul(Ap::fq_from_fp(x), Ap::fq_from_fp(y))); break;
        case 7: r = Ap::fq_to_fp(Ap::fq_sqr(Ap::fq_from_fp(x))); break;
        // the same over Fr (inputs and outputs x 2^256 mod r)
        case 8: r = Ar::fq_to_fp(Ar::fq_mul(Ar::fq_from_fp(x), Ar::fq_from_fp(y))); break;
        case 9: r = Ar::fq_to_fp(Ar::fq_sqr(Ar::fq_from_fp(x))); break;
        default: r = fp_canon(x); break;
    }
    store_fp(out + 2 * i, r);
}

// Points in and out in the 2^256 Montgomery form.  op 0: P + Q (XYZZ add),
// 1: 2P, 2: P + affine(Q.X, Q.Y) with (0,0) as identity -- on the 29-bit
// arithmetic of the MSM kernels; ops 3, 4, 5: the same on the 32-bit
// arithmetic (used by the base generator); ops 6, 7, 8: ops 0, 1, 2 on
// Grumpkin (coordinates in Fr).
__global__ __launch_bounds__(64) void k_test_xyzz(int op, const uint4* __restrict__ p, const uint4* __restrict__ q,
                            uint4* __restrict__ out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Xyzz a = load_xyzz(p, i), b = load_xyzz(q, i);
    if (op >= 3 && op < 6) {
        if (op == 3) xyzz_add(a, b);
        else if (op == 4) a = xyzz_dbl(a);
        else {
            Aff qa;
            qa.x = b.X;
            qa.y = b.Y;
            xyzz_madd(a, qa.x, qa.y, aff_is_id(qa));
        }
        store_xyzz(out, i, a);
        return;
    }
    if (op >= 6) {
        QXyzz qa = Ar::qxyzz_from_xyzz(a), qb = Ar::qxyzz_from_xyzz(b);
        if (op == 6) Ar::qxyzz_add(qa, qb);
        else if (op == 7) qa = Ar::qxyzz_dbl(qa);
        else {
            QAff t;
            t.x = qb.X;
            t.y = qb.Y;
            Ar::qxyzz_madd(qa, t.x, t.y, qaff_is_id(t));
        }
        store_xyzz(out, i, Ar::qxyzz_to_xyzz(qa));
        return;
    }
    QXyzz qa = Ap::qxyzz_from_xyzz(a), qb = Ap::qxyzz_from_xyzz(b);
    if (op == 0) Ap::qxyzz_add(qa, qb);
    else if (op == 1) qa = Ap::qxyzz_dbl(qa);
    else {
        QAff t;
        t.x = qb.X;
        t.y = qb.Y;
        Ap::qxyzz_madd(qa, t.x, t.y, qaff_is_id(t));
    }
    store_xyzz(out, i, Ap::qxyzz_to_xyzz(qa));
}

__global__ __launch_bounds__(64) void k_test_glv(const uint4* __restrict__ scalars, uint4* __restrict__ out,
                           uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 s0 = scalars[2 * i], s1 = scalars[2 * i + 1];
    uint32_t k[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    GlvSplit sp = glv_split(k);
    out[3 * i] = make_uint4(sp.k1[0], sp.k1[1], sp.k1[2], sp.k1[3]);
    out[3 * i + 1] = make_uint4(sp.k2[0], sp.k2[1], sp.k2[2], sp.k2[3]);
    out[3 * i + 2] = make_uint4(sp.neg1, sp.neg2, 0, 0);
}


// ------------------------------------------------------ throughput probes
// Peak rate of the 32x32->64 multiply (v_mad_u64_u32) at full occupancy:
// 8 independent accumulator chains per lane.  mode 0: bare v_mad_u64_u32;
// mode 1: the FIPS accumulate step (mad + carry add); mode 2: field
// multiplies (4 independent chains).  Results are written so nothing is
// dead code.

__global__ __launch_bounds__(256) void k_probe(int mode, uint32_t iters, uint64_t* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 5) {  // 29-bit limbs, latency
        Fq a, b;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            b.v[i] = (t * 2654435761u + i) & ZQ_MASK;
            a.v[i] = (t * 40503u + i * 7u) & ZQ_MASK;
        }
        for (uint32_t it = 0; it < iters; it++) a = Ap::fq_mul(a, b);
        uint64_t x = 0;
#pragma unroll
        for (int i = 0; i < 9; i++) x ^= a.v[i];
        out[t] = x;
        return;
    }
    if (mode == 4) {  // 29-bit limbs, throughput
        Fq a[4], b;
#pragma unroll
        for (int i = 0; i < 9; i++) b.v[i] = (t * 2654435761u + i) & ZQ_MASK;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 9; i++) a[k].v[i] = ((t + k) * 40503u + i * 7u) & ZQ_MASK;
        for (uint32_t it = 0; it < iters; it++)
#pragma unroll
            for (int k = 0; k < 4; k++) a[k] = Ap::fq_mul(a[k], b);
        uint64_t x = 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 9; i++) x ^= a[k].v[i];
        out[t] = x;
        return;
    }
    if (mode == 3) {  // latency: one dependent chain per lane
        Fp a, b;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            b.v[i] = t * 2654435761u + i;
            a.v[i] = t * 40503u + i * 7u;
        }
        for (uint32_t it = 0; it < iters; it++) a = fp_mul(a, b);
        uint64_t x = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) x ^= a.v[i];
        out[t] = x;
        return;
    }
    if (mode == 2) {
        Fp a[4], b;
#pragma unroll
        for (int i = 0; i < 8; i++) b.v[i] = t * 2654435761u + i;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) a[k].v[i] = (t + k) * 40503u + i * 7u;
        for (uint32_t it = 0; it < iters; it++)
#pragma unroll
            for (int k = 0; k < 4; k++) a[k] = fp_mul(a[k], b);
        uint64_t x = 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 8; i++) x ^= a[k].v[i];
        out[t] = x;
        return;
    }
    uint64_t acc[8];
    uint32_t hi[8];
    uint32_t x[8];
    const uint32_t y = t * 2246822519u + 1u;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        acc[i] = t + i;
        hi[i] = 0;
        x[i] = t * 3266489917u + i;
    }
    if (mode == 0) {
        // eight independent multiply-adds in one asm statement: separate
        // statements get a wait state after each from the compiler, which
        // would understate the issue rate
        for (uint32_t it = 0; it < iters; it++) {
            asm volatile(
                "v_mad_u64_u32 %0, s[0:1], %8, %16, %0\n"
                "v_mad_u64_u32 %1, s[0:1], %9, %16, %1\n"
                "v_mad_u64_u32 %2, s[0:1], %10, %16, %2\n"
                "v_mad_u64_u32 %3, s[0:1], %11, %16, %3\n"
                "v_mad_u64_u32 %4, s[0:1], %12, %16, %4\n"
                "v_mad_u64_u32 %5, s[0:1], %13, %16, %5\n"
                "v_mad_u64_u32 %6, s[0:1], %14, %16, %6\n"
                "v_mad_u64_u32 %7, s[0:1], %15, %16, %7\n"
                : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]),
                  "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7])
                : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]),
                  "v"(x[7]), "v"(y)
                : "s0", "s1");
        }
    } else {
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int i = 0; i < 8; i++) mac(acc[i], hi[i], x[i], y);
        }
    }
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r ^= acc[i] + hi[i];
    out[t] = r;
}

// plain 16-byte vector copy, the calibration kernel for HBM byte counters
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                              uint64_t n16) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}



// ------------------------------------------------------------ launchers

static inline unsigned blocks_for(uint64_t n, unsigned bs) {
    return (unsigned)((n + bs - 1) / bs);
}

static inline uint64_t* table_idmask(void* table, uint64_t n, CurveId cv) {
    return reinterpret_cast<uint64_t*>(static_cast<uint8_t*>(table) +
                                       n * table_slots_per_point(cv) * kTableEntryBytes);
}

void launch_import_be(const void* raw, uint64_t n, void* table, uint32_t* bad, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_import_be, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (const uint8_t*)raw, n, (uint4*)table, table_idmask(table, n, CurveId::Bn254),
                       bad);
}

void launch_table_export(CurveId cv, const void* table, uint64_t n, void* pts, hipStream_t st) {
    if (!n) return;
    if (cv == CurveId::Grumpkin)
        hipLaunchKernelGGL((k_table_export<Bn254Fr, 1>), dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const uint4*)table, n, (uint4*)pts);
    else
        hipLaunchKernelGGL((k_table_export<Bn254Fp, 2>), dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const uint4*)table, n, (uint4*)pts);
}

void launch_build_table(CurveId cv, const void* pts, void* table, uint64_t n, hipStream_t st) {
    if (!n) return;
    uint64_t* idm = table_idmask(table, n, cv);
    if (cv == CurveId::Bn254G2)
        hipLaunchKernelGGL(k_build_table_g2, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const uint4*)pts, (uint4*)table, idm, n);
    else if (cv == CurveId::Grumpkin)
        hipLaunchKernelGGL((k_build_table<Bn254Fr, 1>), dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const uint4*)pts, (uint4*)table, idm, n);
    else
        hipLaunchKernelGGL((k_build_table<Bn254Fp, 2>), dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const uint4*)pts, (uint4*)table, idm, n);
}

void launch_shift_windows(void* table, uint64_t n, uint64_t stride, uint32_t c, uint32_t nwin,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_shift_windows, dim3(blocks_for(n, 256)), dim3(256), 0, st, (uint4*)table, n,
                       stride, c, nwin);
}

void launch_decompose(CurveId cv, const MsmDesc* msms, const uint64_t* msm_dtile0, uint32_t nmsm,
                      uint32_t dtiles, const SegDesc* segs, const void* scalars,
                      const void* table, uint64_t table_n, void* digits, uint32_t* hist,
                      uint32_t* err, hipStream_t st) {
    if (!dtiles) return;
    const uint64_t* idm = table_idmask(const_cast<void*>(table), table_n, cv);
    if (cv == CurveId::Grumpkin)
        hipLaunchKernelGGL((k_decompose<1, false>), dim3(dtiles), dim3(kDecThreads), 0, st, msms, msm_dtile0,
                           nmsm, segs, (const uint4*)scalars, idm, (uint2*)digits, hist, err);
    else if (cv == CurveId::Bn254G2)
        hipLaunchKernelGGL((k_decompose<1, true>), dim3(dtiles), dim3(kDecThreads), 0, st, msms, msm_dtile0,
                           nmsm, segs, (const uint4*)scalars, idm, (uint2*)digits, hist, err);
    else
        hipLaunchKernelGGL((k_decompose<2, false>), dim3(dtiles), dim3(kDecThreads), 0, st, msms, msm_dtile0,
                           nmsm, segs, (const uint4*)scalars, idm, (uint2*)digits, hist, err);
}

void launch_sort_pass(const SegDesc* segs, uint32_t seg_lo, uint32_t nseg, const uint32_t* tile_seg,
                      uint32_t tile_lo, uint32_t ntiles, const uint32_t* chunk_seg, uint32_t chunk_lo,
                      uint32_t nchunks, const void* in_keys, const uint32_t* in_vals, int pass,
                      uint32_t* hist, uint32_t* part, uint32_t* seg_valid, uint16_t* out_keys,
                      uint32_t* out_vals, hipStream_t st, bool count) {
    if (!ntiles) return;
    if (count)  // else the tile histograms are already in hist
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(kSortThreads), 0, st, segs, tile_seg, tile_lo,
                           in_keys, pass, (const uint32_t*)seg_valid, hist);
    hipLaunchKernelGGL(k_scan_up, dim3(nchunks), dim3(256), 0, st, segs, chunk_seg, chunk_lo,
                       (const uint32_t*)hist, part);
    hipLaunchKernelGGL(k_scan_spine, dim3(nseg), dim3(256), 0, st, segs, seg_lo, part);
    hipLaunchKernelGGL(k_scan_down, dim3(nchunks), dim3(256), 0, st, segs, chunk_seg, chunk_lo, hist,
                       (const uint32_t*)part, pass, seg_valid);
    hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(kSortThreads), 0, st, segs, tile_seg, tile_lo,
                       in_keys, in_vals, pass, (const uint32_t*)hist, (const uint32_t*)seg_valid,
                       out_keys, out_vals);
}

void launch_fill_zero(void* p, uint64_t bytes, hipStream_t st) {
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    unsigned blocks = blocks_for(n16, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, st, (uint4*)p, n16);
}

uint64_t accumulate_capacity(int device, CurveId cv) {
    int cus = 0, blocks = 0;
    const hipError_t eo = cv == CurveId::Bn254G2
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_accumulate<G2Curve>, 256, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_accumulate<G1Curve<Bn254Fp>>, 256, 0);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        eo != hipSuccess || cus <= 0 || blocks <= 0) {
        (void)hipGetLastError();
        return 1u << 17;
    }
    return (uint64_t)cus * (uint64_t)blocks * 256u;
}

void launch_lane_plan(const SegDesc* segs, const uint32_t* seg_valid, uint32_t seg_lo, uint32_t nseg,
                      uint64_t lane_lo, uint64_t lane_hi, uint64_t* seg_lane0, uint32_t* T, hipStream_t st) {
    hipLaunchKernelGGL(k_lane_plan, dim3(1), dim3(256), 0, st, segs, seg_valid, seg_lo, nseg, lane_lo,
                       lane_hi - lane_lo, seg_lane0, T);
}

void launch_accumulate(CurveId cv, const SegDesc* segs, const uint64_t* seg_lane0, uint32_t seg_lo,
                       uint32_t nseg, const MsmDesc* msms, const uint32_t* seg_valid, const uint16_t* keys,
                       const uint32_t* vals, const void* table, const uint32_t* T, uint64_t lane_lo,
                       uint64_t lane_hi, void* buckets, uint32_t* rec_keys, void* rec_pts, hipStream_t st) {
    if (lane_hi <= lane_lo) return;
#define ZK_ACC(C)                                                                                  \
    hipLaunchKernelGGL(k_accumulate<C>, dim3(blocks_for(lane_hi - lane_lo, 256)), dim3(256), 0, st, \
                       segs, seg_lane0, seg_lo, nseg, msms, seg_valid, keys, vals, (const uint4*)table, \
                       T, lane_lo, lane_hi, (uint4*)buckets, rec_keys, (uint4*)rec_pts)
    if (cv == CurveId::Grumpkin) ZK_ACC(G1Curve<Bn254Fr>);
    else if (cv == CurveId::Bn254G2) ZK_ACC(G2Curve);
    else ZK_ACC(G1Curve<Bn254Fp>);
#undef ZK_ACC
}

template <class C>
static void merge_runs_t(const uint32_t* keys, const void* pts, uint64_t nrec, uint32_t limit,
                         void* dest, uint32_t* counts, uint32_t* lists, uint32_t list_cap,
                         hipStream_t st) {
    uint32_t* long_list = lists;
    uint32_t* mid_list = lists + list_cap;
    hipLaunchKernelGGL(k_merge_runs<C>, dim3(blocks_for((nrec + 1) / 2, 256)), dim3(256), 0, st, keys,
                       (const uint4*)pts, nrec, limit, (uint4*)dest, counts, long_list,
                       counts + 1, mid_list, list_cap);
    // grid-stride over the lists (their lengths are known on the device only)
    hipLaunchKernelGGL(k_merge_mid<C>, dim3(std::min<uint32_t>(blocks_for(list_cap, 32), 1024)), dim3(256),
                       0, st, keys, (const uint4*)pts, nrec, (const uint32_t*)(counts + 1),
                       (const uint32_t*)mid_list, list_cap, (uint4*)dest);
    hipLaunchKernelGGL(k_merge_long<C>, dim3(std::min<uint32_t>(list_cap, 2048)), dim3(256), 0, st, keys,
                       (const uint4*)pts, nrec, (const uint32_t*)counts,
                       (const uint32_t*)long_list, list_cap, (uint4*)dest);
}

void launch_merge_runs(CurveId cv, const uint32_t* keys, const void* pts, uint64_t nrec,
                       uint32_t limit, void* dest, uint32_t* counts, uint32_t* lists,
                       uint32_t list_cap, hipStream_t st) {
    if (!nrec || !list_cap) return;
    if (cv == CurveId::Grumpkin)
        merge_runs_t<G1Curve<Bn254Fr>>(keys, pts, nrec, limit, dest, counts, lists, list_cap, st);
    else if (cv == CurveId::Bn254G2)
        merge_runs_t<G2Curve>(keys, pts, nrec, limit, dest, counts, lists, list_cap, st);
    else
        merge_runs_t<G1Curve<Bn254Fp>>(keys, pts, nrec, limit, dest, counts, lists, list_cap, st);
}

template <class C>
static void bucket_reduce_t(const SegDesc* segs, const uint64_t* seg_wave0, const uint64_t* seg_rc0,
                            uint32_t nseg, uint32_t seg_lo, uint32_t seg_hi, uint64_t wave_lo,
                            uint64_t wave_hi, uint32_t per, const void* buckets, void* rc, void* bits,
                            bool host_form, hipStream_t st) {
    hipLaunchKernelGGL(k_br_rowcol<C>, dim3(blocks_for(wave_hi - wave_lo, 4)), dim3(256), 0, st, segs,
                       seg_wave0, seg_rc0, nseg, wave_lo, wave_hi, per, (const uint4*)buckets, (uint4*)rc);
    const uint64_t bit_items = (uint64_t)(seg_hi - seg_lo) * 16;
    hipLaunchKernelGGL(k_br_bits<C>, dim3(blocks_for(bit_items, 4)), dim3(256), 0, st, segs,
                       seg_rc0, seg_lo, seg_hi, (const uint4*)rc, (uint4*)bits, host_form ? 1 : 0);
}

void launch_bucket_reduce(CurveId cv, const SegDesc* segs, const uint64_t* seg_wave0,
                          const uint64_t* seg_rc0, uint32_t nseg, uint32_t seg_lo, uint32_t seg_hi,
                          uint64_t wave_lo, uint64_t wave_hi, uint32_t per, const void* buckets, void* rc,
                          void* bits, bool host_form, hipStream_t st) {
    if (wave_hi <= wave_lo || seg_hi <= seg_lo) return;
    if (cv == CurveId::Grumpkin)
        bucket_reduce_t<G1Curve<Bn254Fr>>(segs, seg_wave0, seg_rc0, nseg, seg_lo, seg_hi, wave_lo, wave_hi, per,
                                          buckets, rc, bits, host_form, st);
    else if (cv == CurveId::Bn254G2)
        bucket_reduce_t<G2Curve>(segs, seg_wave0, seg_rc0, nseg, seg_lo, seg_hi, wave_lo, wave_hi, per, buckets,
                                 rc, bits, true, st);
    else
        bucket_reduce_t<G1Curve<Bn254Fp>>(segs, seg_wave0, seg_rc0, nseg, seg_lo, seg_hi, wave_lo, wave_hi, per,
                                          buckets, rc, bits, host_form, st);
}

void launch_tail(CurveId cv, const SegDesc* segs, const MsmDesc* msms, uint32_t nseg, uint32_t nmsm,
                 const void* bits, void* wins, void* out, hipStream_t st) {
    if (!nseg || !nmsm) return;
    if (cv == CurveId::Grumpkin) {
        hipLaunchKernelGGL(k_tail_windows<G1Curve<Bn254Fr>>, dim3((nseg + 3) / 4), dim3(64), 0, st, segs,
                           msms, nseg, (const uint4*)bits, (uint4*)wins);
        hipLaunchKernelGGL(k_tail_combine<G1Curve<Bn254Fr>>, dim3((nmsm + 63) / 64), dim3(64), 0, st, msms,
                           nmsm, (const uint4*)wins, (uint4*)out);
    } else if (cv == CurveId::Bn254) {
        hipLaunchKernelGGL(k_tail_windows<G1Curve<Bn254Fp>>, dim3((nseg + 3) / 4), dim3(64), 0, st, segs,
                           msms, nseg, (const uint4*)bits, (uint4*)wins);
        hipLaunchKernelGGL(k_tail_combine<G1Curve<Bn254Fp>>, dim3((nmsm + 63) / 64), dim3(64), 0, st, msms,
                           nmsm, (const uint4*)wins, (uint4*)out);
    }
}

void launch_gen_bases(const void* ks, const void* gtab, uint64_t n, void* table,
                      hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_gen_bases, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (const uint4*)ks, (const uint4*)gtab, n, (uint4*)table,
                       table_idmask(table, n, CurveId::Bn254));
}

void launch_test_fp(int op, const void* a, const void* b, void* out, uint64_t n,
                    hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_test_fp, dim3(blocks_for(n, 64)), dim3(64), 0, st, op,
                       (const uint4*)a, (const uint4*)b, (uint4*)out, n);
}

void launch_test_xyzz(int op, const void* p, const void* q, void* out, uint64_t n,
                      hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_test_xyzz, dim3(blocks_for(n, 64)), dim3(64), 0, st, op,
                       (const uint4*)p, (const uint4*)q, (uint4*)out, n);
}

void launch_gen_points_g2(const void* ks, const void* gtab, uint64_t n, void* out, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_gen_points_g2, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (const uint4*)ks, (const uint4*)gtab, n, (uint4*)out);
}

void launch_test_g2(int op, const void* p, const void* q, void* out, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_test_g2, dim3(blocks_for(n, 64)), dim3(64), 0, st, op,
                       (const uint4*)p, (const uint4*)q, (uint4*)out, n);
}

void launch_gen_points_grumpkin(const void* ks, const void* gtab, uint64_t n, void* out,
                                hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_gen_points<Bn254Fr>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (const uint4*)ks, (const uint4*)gtab, n, (uint4*)out);
}

void launch_test_glv(const void* scalars, void* out, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_test_glv, dim3(blocks_for(n, 64)), dim3(64), 0, st,
                       (const uint4*)scalars, (uint4*)out, n);
}

void launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t st) {
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    unsigned blocks = (unsigned)std::min<uint64_t>((n16 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n16);
}

void launch_probe(int mode, uint32_t blocks, uint32_t iters, uint64_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_probe, dim3(blocks), dim3(256), 0, st, mode, iters, out);
}

}  // namespace zk
