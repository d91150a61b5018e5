// HIP kernels of the BN254 G1 MSM for gfx950 (wave64).
//
// Pipeline (per batch of MSMs, all on one stream):
//   k_build_table     bases -> [P, phi(P)] table              (glv_kernels.metal:260-297)
//   k_decompose       scalars -> GLV halves -> signed digits  (glv_kernels.metal:172-256,
//                                                              msm_kernels.metal:399-439)
//   k_sort_hist/scan/scatter x2   LSD radix sort of bucket keys per segment
//                                 (replaces MSMEngine.swift:595-661, CPU counting sort)
//   k_accumulate      load-balanced XYZZ mixed additions into buckets
//                                 (replaces msm_kernels.metal:16-62)
//   k_merge           segmented reduction of bucket partials across lanes
//   k_bucket_reduce   per-group weighted bucket sums      (replaces msm_kernels.metal:152-248)
//   k_merge           -> one XYZZ window sum per segment
// The window combine (Horner over windows) and affine normalisation are a
// ~1000-multiply serial tail and run on the host (MSMEngine.swift:899-904
// does the same).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "curve.hpp"
#include "glv.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace zk {

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

__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << __lane_id()) - 1ull;
}

// mask of active lanes whose 8-bit digit equals this lane's
__device__ __forceinline__ uint64_t match_digit8(uint32_t digit, bool active) {
    uint64_t peers = __ballot(active);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        bool bit = (digit >> b) & 1u;
        uint64_t m = __ballot(active && bit);
        peers &= bit ? m : ~m;
    }
    return active ? peers : 0ull;
}

// -------------------------------------------------------- base table / phi

__global__ __launch_bounds__(256) void k_build_table(const uint4* __restrict__ pts,
                                                     uint4* __restrict__ table, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Aff a = load_aff(pts, i);
    a.x = fp_canon(a.x);
    a.y = fp_canon(a.y);
    store_aff(table, 2 * i, a);
    Aff e;
    e.x = fp_mul(fp_beta(), a.x);  // (0,0) maps to (0,0)
    e.y = a.y;
    store_aff(table, 2 * i + 1, e);
}

// --------------------------------------------------- GLV + signed digits

__global__ __launch_bounds__(256) void k_decompose(const MsmDesc* __restrict__ msms,
                                                   const uint64_t* __restrict__ msm_scalar0,
                                                   uint32_t nmsm, const SegDesc* __restrict__ segs,
                                                   const uint4* __restrict__ scalars,
                                                   uint64_t total, uint2* __restrict__ digits,
                                                   uint32_t* __restrict__ err) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    uint32_t j = upper_index(msm_scalar0, nmsm, g);
    const MsmDesc m = msms[j];
    uint64_t i = g - m.scalar0;
    uint4 s0 = scalars[2 * g], s1 = scalars[2 * g + 1];
    uint32_t k[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    GlvSplit sp = glv_split(k);
    const uint32_t c = m.c, half = 1u << (c - 1), full = 1u << c;
    uint32_t carry1 = 0, carry2 = 0;
    for (uint32_t w = 0; w < m.nwin; w++) {
        uint32_t d1 = window_bits128(sp.k1, w, c) + carry1;
        uint32_t d2 = window_bits128(sp.k2, w, c) + carry2;
        uint32_t neg1 = sp.neg1 ? 1u : 0u, neg2 = sp.neg2 ? 1u : 0u;
        carry1 = d1 > half;
        carry2 = d2 > half;
        if (carry1) { d1 = full - d1; neg1 ^= 1u; }
        if (carry2) { d2 = full - d2; neg2 ^= 1u; }
        uint32_t w1 = d1 ? ((d1 - 1u) | (neg1 << 31)) : ZK_KEY_NONE;
        uint32_t w2 = d2 ? ((d2 - 1u) | (neg2 << 31)) : ZK_KEY_NONE;
        const uint64_t e = segs[m.seg0 + w].ent0 + 2 * i;
        digits[e >> 1] = make_uint2(w1, w2);
    }
    if (carry1 | carry2) atomicOr(err, 1u);
}

// ------------------------------------------------------ radix sort passes
// Two stable LSD passes of 8 bits over the 16-bit bucket key, independently
// per segment.  Tiles of 4096 entries; per tile 256-bin histograms are
// written bin-major per segment (hist[tile0*256 + bin*ntiles + t]) so one
// exclusive scan per segment yields every (bin, tile) destination.

__device__ __forceinline__ uint32_t sort_key(const void* in, int pass, uint64_t pos) {
    if (pass == 0) return reinterpret_cast<const uint32_t*>(in)[pos] & 0xffffu;
    return reinterpret_cast<const uint16_t*>(in)[pos];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const SegDesc* __restrict__ segs,
                                                            const uint32_t* __restrict__ tile_seg,
                                                            const void* __restrict__ in, int pass,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    const uint32_t t = blockIdx.x;
    const SegDesc sg = segs[tile_seg[t]];
    const uint32_t tin = t - sg.tile0;
    const uint32_t start = tin * kSortTile;
    const uint32_t len = min((uint32_t)kSortTile, sg.len - start);
    const uint32_t shift = pass ? 8u : 0u;
    cnt[threadIdx.x] = 0;
    __syncthreads();
    for (int r = 0; r < kSortItems; r++) {
        uint32_t idx = r * kSortThreads + threadIdx.x;
        bool act = idx < len;
        uint32_t digit = act ? (sort_key(in, pass, sg.ent0 + start + idx) >> shift) & 0xffu : 0u;
        uint64_t peers = match_digit8(digit, act);
        if (act && (peers & lanemask_lt()) == 0) atomicAdd(&cnt[digit], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[(uint64_t)sg.tile0 * 256 + (uint64_t)threadIdx.x * sg.ntiles + tin] = cnt[threadIdx.x];
}

// One block per segment: exclusive scan of its 256*ntiles counts, turned into
// global destinations (+ ent0).  Pass 1 also records the number of entries
// with a real bucket (everything before bin 0xFF, which only the "no bucket"
// key 0xFFFF reaches).
__global__ __launch_bounds__(256) void k_sort_scan(const SegDesc* __restrict__ segs,
                                                   uint32_t* __restrict__ hist, int pass,
                                                   uint32_t* __restrict__ seg_valid) {
    __shared__ uint32_t part[256];
    const SegDesc sg = segs[blockIdx.x];
    uint32_t* h = hist + (uint64_t)sg.tile0 * 256;
    const uint32_t M = 256u * sg.ntiles;
    const uint32_t chunk = (M + 255u) / 256u;
    const uint32_t b = min(M, threadIdx.x * chunk), e = min(M, b + chunk);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += h[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < 256; i++) {
            uint32_t v = part[i];
            part[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    uint32_t acc = part[threadIdx.x];
    const uint32_t none_bin = 0xffu * sg.ntiles;
    for (uint32_t i = b; i < e; i++) {
        uint32_t v = h[i];
        if (pass == 1 && i == none_bin) seg_valid[blockIdx.x] = acc;
        h[i] = (uint32_t)sg.ent0 + acc;
        acc += v;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const SegDesc* __restrict__ segs, const uint32_t* __restrict__ tile_seg,
    const void* __restrict__ in_keys, const uint32_t* __restrict__ in_vals, int pass,
    const uint32_t* __restrict__ hist, uint16_t* __restrict__ out_keys,
    uint32_t* __restrict__ out_vals) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[4][256];
    const uint32_t t = blockIdx.x;
    const SegDesc sg = segs[tile_seg[t]];
    const uint32_t tin = t - sg.tile0;
    const uint32_t start = tin * kSortTile;
    const uint32_t len = min((uint32_t)kSortTile, sg.len - start);
    const uint32_t shift = pass ? 8u : 0u;
    const uint32_t wave = threadIdx.x >> 6;
    run[threadIdx.x] = hist[(uint64_t)sg.tile0 * 256 + (uint64_t)threadIdx.x * sg.ntiles + tin];
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t idx = r * kSortThreads + threadIdx.x;
        const bool act = idx < len;
        const uint64_t pos = sg.ent0 + start + idx;
        uint32_t key = 0, val = 0;
        if (act) {
            if (pass == 0) {
                uint32_t wd = reinterpret_cast<const uint32_t*>(in_keys)[pos];
                key = wd & 0xffffu;
                val = (start + idx) | (wd & 0x80000000u);
            } else {
                key = reinterpret_cast<const uint16_t*>(in_keys)[pos];
                val = in_vals[pos];
            }
        }
        const uint32_t digit = (key >> shift) & 0xffu;
#pragma unroll
        for (int w = 0; w < 4; w++) wc[w][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t peers = match_digit8(digit, act);
        const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (act && rank == 0) wc[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            const uint32_t d = threadIdx.x;
            uint32_t base = run[d];
            uint32_t c0 = wc[0][d], c1 = wc[1][d], c2 = wc[2][d], c3 = wc[3][d];
            wc[0][d] = base;
            wc[1][d] = base + c0;
            wc[2][d] = base + c0 + c1;
            wc[3][d] = base + c0 + c1 + c2;
            run[d] = base + c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (act) {
            const uint32_t dst = wc[wave][digit] + rank;
            if (pass == 0 || key != ZK_KEY_NONE) {
                out_keys[dst] = (uint16_t)key;
                out_vals[dst] = val;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ utilities

__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ p, uint64_t n16) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

// --------------------------------------------------- bucket accumulation
// Lane L of segment s owns sorted entries [ent0 + l*T, ent0 + (l+1)*T) clipped
// to the segment's valid entries.  It walks them once with mixed additions;
// a bucket run that starts and ends inside the range is final and is written
// to the bucket array, the (at most two) runs that cross the range ends go to
// merge records 2L (first run) and 2L+1 (last run).

__device__ __forceinline__ void put_record(uint32_t* keys, uint4* pts, uint64_t slot,
                                           uint32_t key, bool valid, const Xyzz& v) {
    keys[slot] = key | (valid ? 0u : ZK_REC_INVALID);
    if (valid) store_xyzz(pts, slot, v);
}

__global__ __launch_bounds__(256) void k_accumulate(
    const SegDesc* __restrict__ segs, const uint64_t* __restrict__ seg_lane0, uint32_t nseg,
    const MsmDesc* __restrict__ msms, const uint32_t* __restrict__ seg_valid,
    const uint16_t* __restrict__ keys, const uint32_t* __restrict__ vals,
    const uint4* __restrict__ table, uint32_t T, uint64_t total_lanes, uint4* __restrict__ buckets,
    uint32_t* __restrict__ rec_keys, uint4* __restrict__ rec_pts) {
    const uint64_t L = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= total_lanes) return;
    const uint32_t s = upper_index(seg_lane0, nseg, L);
    const SegDesc sg = segs[s];
    const uint64_t l = L - sg.lane0;
    const uint64_t segend = sg.ent0 + seg_valid[s];
    const uint64_t b = sg.ent0 + l * T;
    const uint64_t e = min(b + T, segend);
    if (b >= e) {
        rec_keys[2 * L] = ZK_REC_INVALID | ZK_REC_KEYMASK;
        rec_keys[2 * L + 1] = ZK_REC_INVALID | ZK_REC_KEYMASK;
        return;
    }
    const uint4* tb = table + msms[sg.msm].table0 * 4;
    uint32_t cur = keys[b];
    const bool head_open = (b > sg.ent0) && keys[b - 1] == cur;
    bool first = true;
    Xyzz acc = xyzz_identity();
    for (uint64_t p = b; p < e; p++) {
        const uint32_t k = keys[p];
        const uint32_t v = vals[p];
        if (k != cur) {
            // run `cur` ended at p-1 and is not the lane's last run
            const uint32_t gk = (uint32_t)(sg.bkt0 + cur);
            if (first && head_open) {
                put_record(rec_keys, rec_pts, 2 * L, gk, true, acc);
            } else {
                store_xyzz(buckets, gk, acc);
                if (first) put_record(rec_keys, rec_pts, 2 * L, gk, false, acc);
            }
            first = false;
            acc = xyzz_identity();
            cur = k;
        }
        Aff q = load_aff(tb, v & 0x7fffffffu);
        const bool q_id = aff_is_id(q);
        if (v >> 31) q.y = fp_neg(q.y);
        xyzz_madd(acc, q.x, q.y, q_id);
    }
    const bool tail_open = (e < segend) && keys[e] == cur;
    const uint32_t gk = (uint32_t)(sg.bkt0 + cur);
    if (first) {
        const bool open = head_open || tail_open;
        if (!open) store_xyzz(buckets, gk, acc);
        put_record(rec_keys, rec_pts, 2 * L, gk, open, acc);
        put_record(rec_keys, rec_pts, 2 * L + 1, gk, false, acc);
    } else {
        if (!tail_open) store_xyzz(buckets, gk, acc);
        put_record(rec_keys, rec_pts, 2 * L + 1, gk, tail_open, acc);
    }
}

// ------------------------------------------------- segmented merge level
// Records [L*T, (L+1)*T) per lane; runs of equal key are summed (padding
// records skipped).  A run closed on both sides with at least one payload is
// final and written to dest[key]; open runs become the lane's output records.

__global__ __launch_bounds__(256) void k_merge(const uint32_t* __restrict__ in_keys,
                                               const uint4* __restrict__ in_pts, uint64_t nrec,
                                               uint32_t T, uint64_t lanes,
                                               uint32_t* __restrict__ out_keys,
                                               uint4* __restrict__ out_pts,
                                               uint4* __restrict__ dest) {
    const uint64_t L = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= lanes) return;
    const uint64_t b = L * T;
    const uint64_t e = min(b + T, nrec);
    if (b >= e) {
        out_keys[2 * L] = ZK_REC_INVALID | ZK_REC_KEYMASK;
        out_keys[2 * L + 1] = ZK_REC_INVALID | ZK_REC_KEYMASK;
        return;
    }
    uint32_t cur = in_keys[b] & ZK_REC_KEYMASK;
    const bool head_open = (b > 0) && (in_keys[b - 1] & ZK_REC_KEYMASK) == cur;
    bool first = true, has = false;
    Xyzz acc = xyzz_identity();
    for (uint64_t p = b; p < e; p++) {
        const uint32_t kr = in_keys[p];
        const uint32_t k = kr & ZK_REC_KEYMASK;
        if (k != cur) {
            if (first && head_open) {
                put_record(out_keys, out_pts, 2 * L, cur, has, acc);
            } else {
                if (has) store_xyzz(dest, cur, acc);
                if (first) put_record(out_keys, out_pts, 2 * L, cur, false, acc);
            }
            first = false;
            has = false;
            acc = xyzz_identity();
            cur = k;
        }
        if (!(kr & ZK_REC_INVALID)) {
            Xyzz q = load_xyzz(in_pts, p);
            if (has) xyzz_add(acc, q);
            else acc = q;
            has = true;
        }
    }
    const bool tail_open = (e < nrec) && (in_keys[e] & ZK_REC_KEYMASK) == cur;
    if (first) {
        const bool open = head_open || tail_open;
        if (!open && has) store_xyzz(dest, cur, acc);
        put_record(out_keys, out_pts, 2 * L, cur, open && has, acc);
        put_record(out_keys, out_pts, 2 * L + 1, cur, false, acc);
    } else {
        if (!tail_open && has) store_xyzz(dest, cur, acc);
        put_record(out_keys, out_pts, 2 * L + 1, cur, tail_open && has, acc);
    }
}

// ------------------------------------------------------ bucket reduction
// Group g of segment s covers buckets [lo, lo+G): with the running sum from
// the top, acc = sum (k - lo + 1) B_k; adding lo * run gives
// sum (k + 1) B_k, the group's share of the window sum (bucket key k holds
// digit magnitude k + 1).  Groups become records keyed by segment.

__device__ __forceinline__ Xyzz xyzz_mul_small(const Xyzz& p, uint32_t m) {
    Xyzz r = xyzz_identity();
    if (m == 0) return r;
    int top = 31 - __clz(m);
    r = p;
    for (int bit = top - 1; bit >= 0; bit--) {
        r = xyzz_dbl(r);
        if ((m >> bit) & 1u) xyzz_add(r, p);
    }
    return r;
}

__global__ __launch_bounds__(256) void k_bucket_reduce(
    const SegDesc* __restrict__ segs, const uint64_t* __restrict__ seg_grp0, uint32_t nseg,
    uint32_t G, uint64_t total_groups, const uint4* __restrict__ buckets,
    uint32_t* __restrict__ rec_keys, uint4* __restrict__ rec_pts) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total_groups) return;
    const uint32_t s = upper_index(seg_grp0, nseg, gid);
    const SegDesc sg = segs[s];
    const uint32_t g = (uint32_t)(gid - sg.grp0);
    const uint32_t lo = g * G;
    const uint32_t hi = min(lo + G, sg.nbk);
    Xyzz run = xyzz_identity(), acc = xyzz_identity();
    for (uint32_t k = hi; k > lo; k--) {
        Xyzz bk = load_xyzz(buckets, sg.bkt0 + k - 1);
        xyzz_add(run, bk);
        xyzz_add(acc, run);
    }
    if (lo > 0) {
        Xyzz t = xyzz_mul_small(run, lo);
        xyzz_add(acc, t);
    }
    rec_keys[gid] = s;
    store_xyzz(rec_pts, gid, acc);
}

// ------------------------------------------------ synthetic base table
// table[t0 + 2i] = k_i * G, table[t0 + 2i + 1] = phi(k_i * G), from a host
// table gtab[w*256 + b] = b * 2^(8w) * G (affine, entry 0 = (0,0)).

__global__ __launch_bounds__(256) void k_gen_bases(const uint4* __restrict__ ks,
                                                   const uint4* __restrict__ gtab, uint64_t n,
                                                   uint4* __restrict__ table) {
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
    store_aff(table, 2 * i, a);
    Aff e;
    e.x = fp_mul(fp_beta(), a.x);
    e.y = a.y;
    store_aff(table, 2 * i + 1, e);
}

// ------------------------------------------------------ test entry points
// Field / curve primitives exposed for the GPU unit tests.

__global__ void k_test_fp(int op, const uint4* __restrict__ a, const uint4* __restrict__ b,
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
        default: r = fp_canon(x); break;
    }
    store_fp(out + 2 * i, r);
}

// op 0: P + Q (XYZZ add), 1: 2P, 2: P + affine(Q.X, Q.Y) with (0,0) as identity
__global__ void k_test_xyzz(int op, const uint4* __restrict__ p, const uint4* __restrict__ q,
                            uint4* __restrict__ out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Xyzz a = load_xyzz(p, i), b = load_xyzz(q, i);
    if (op == 0) xyzz_add(a, b);
    else if (op == 1) a = xyzz_dbl(a);
    else {
        Aff qa;
        qa.x = b.X;
        qa.y = b.Y;
        xyzz_madd(a, qa.x, qa.y, aff_is_id(qa));
    }
    store_xyzz(out, i, a);
}

__global__ void k_test_glv(const uint4* __restrict__ scalars, uint4* __restrict__ out,
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
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                uint64_t c;
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[i]), "=s"(c) : "v"(x[i]), "v"(y));
            }
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

void launch_probe(int mode, uint32_t blocks, uint32_t iters, uint64_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_probe, dim3(blocks), dim3(256), 0, st, mode, iters, out);
}

// ------------------------------------------------------------ launchers

static inline unsigned blocks_for(uint64_t n, unsigned bs) {
    return (unsigned)((n + bs - 1) / bs);
}

void launch_build_table(const void* pts, void* table, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_build_table, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (const uint4*)pts, (uint4*)table, n);
}

void launch_decompose(const MsmDesc* msms, const uint64_t* msm_scalar0, uint32_t nmsm,
                      const SegDesc* segs, const void* scalars, uint64_t total, void* digits,
                      uint32_t* err, hipStream_t st) {
    if (!total) return;
    hipLaunchKernelGGL(k_decompose, dim3(blocks_for(total, 256)), dim3(256), 0, st, msms,
                       msm_scalar0, nmsm, segs, (const uint4*)scalars, total, (uint2*)digits, err);
}

void launch_sort_pass(const SegDesc* segs, uint32_t nseg, const uint32_t* tile_seg,
                      uint32_t ntiles, const void* in_keys, const uint32_t* in_vals, int pass,
                      uint32_t* hist, uint32_t* seg_valid, uint16_t* out_keys,
                      uint32_t* out_vals, hipStream_t st) {
    if (!ntiles) return;
    hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(kSortThreads), 0, st, segs, tile_seg,
                       in_keys, pass, hist);
    hipLaunchKernelGGL(k_sort_scan, dim3(nseg), dim3(256), 0, st, segs, hist, pass, seg_valid);
    hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(kSortThreads), 0, st, segs, tile_seg,
                       in_keys, in_vals, pass, (const uint32_t*)hist, out_keys, out_vals);
}

void launch_fill_zero(void* p, uint64_t bytes, hipStream_t st) {
    uint64_t n16 = bytes / 16;
    if (!n16) return;
    unsigned blocks = blocks_for(n16, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, st, (uint4*)p, n16);
}

void launch_accumulate(const SegDesc* segs, const uint64_t* seg_lane0, uint32_t nseg,
                       const MsmDesc* msms, const uint32_t* seg_valid, const uint16_t* keys,
                       const uint32_t* vals, const void* table, uint32_t T, uint64_t lanes,
                       void* buckets, uint32_t* rec_keys, void* rec_pts, hipStream_t st) {
    if (!lanes) return;
    hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(lanes, 256)), dim3(256), 0, st, segs,
                       seg_lane0, nseg, msms, seg_valid, keys, vals, (const uint4*)table, T,
                       lanes, (uint4*)buckets, rec_keys, (uint4*)rec_pts);
}

void launch_merge(const uint32_t* in_keys, const void* in_pts, uint64_t nrec, uint32_t T,
                  uint32_t* out_keys, void* out_pts, void* dest, hipStream_t st) {
    uint64_t lanes = (nrec + T - 1) / T;
    if (!lanes) return;
    hipLaunchKernelGGL(k_merge, dim3(blocks_for(lanes, 256)), dim3(256), 0, st, in_keys,
                       (const uint4*)in_pts, nrec, T, lanes, out_keys, (uint4*)out_pts,
                       (uint4*)dest);
}

void launch_bucket_reduce(const SegDesc* segs, const uint64_t* seg_grp0, uint32_t nseg,
                          uint32_t G, uint64_t groups, const void* buckets, uint32_t* rec_keys,
                          void* rec_pts, hipStream_t st) {
    if (!groups) return;
    hipLaunchKernelGGL(k_bucket_reduce, dim3(blocks_for(groups, 256)), dim3(256), 0, st, segs,
                       seg_grp0, nseg, G, groups, (const uint4*)buckets, rec_keys,
                       (uint4*)rec_pts);
}

void launch_gen_bases(const void* ks, const void* gtab, uint64_t n, void* table,
                      hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_gen_bases, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (const uint4*)ks, (const uint4*)gtab, n, (uint4*)table);
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

void launch_test_glv(const void* scalars, void* out, uint64_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_test_glv, dim3(blocks_for(n, 64)), dim3(64), 0, st,
                       (const uint4*)scalars, (uint4*)out, n);
}

}  // namespace zk
