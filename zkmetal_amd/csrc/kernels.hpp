// Host-callable launchers for the MSM kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plan.hpp"

namespace zk {

// A table of n points of curve cv takes table_bytes(n, cv): n rows of
// table_slots_per_point(cv) 128-byte slots, then a bitmap with bit i set when
// point i is the identity (written by the table builders, read by the
// decomposition).
inline uint64_t table_bytes(uint64_t n, CurveId cv = CurveId::Bn254) {
    return n * table_slots_per_point(cv) * kTableEntryBytes + (n + 63) / 64 * 8;
}
// BN254 G1 only: SRS import from big-endian standard-form coordinates
void launch_import_be(const void* raw, uint64_t n, void* table, uint32_t* bad, hipStream_t st);
void launch_table_export(CurveId cv, const void* table, uint64_t n, void* pts, hipStream_t st);
void launch_build_table(CurveId cv, const void* pts, void* table, uint64_t n, hipStream_t st);
// BN254 G1 table of n bases: window-shifted copies 2^(c w) P_i, phi(2^(c w) P_i)
// at entries w * stride + 2i, 2i + 1 for w = 1 .. nwin-1 (the table's
// allocation holds nwin * stride entries; stride * 128 >= table_bytes(n))
void launch_shift_windows(void* table, uint64_t n, uint64_t stride, uint32_t c, uint32_t nwin,
                          hipStream_t st);
// one block per first-pass sort tile of each MSM (msm_dtile0, dtiles blocks):
// digits, and the first sort pass's tile histograms into hist
void launch_decompose(CurveId cv, const MsmDesc* msms, const uint64_t* msm_dtile0, uint32_t nmsm,
                      uint32_t dtiles, const SegDesc* segs, const void* scalars,
                      const void* table, uint64_t table_n, void* digits, uint32_t* hist,
                      uint32_t* err, hipStream_t st);
// one radix pass over segments [seg_lo, seg_lo + nseg): their tiles
// [tile_lo, tile_lo + ntiles) and scan chunks [chunk_lo, chunk_lo + nchunks)
void launch_sort_pass(const SegDesc* segs, uint32_t seg_lo, uint32_t nseg, const uint32_t* tile_seg,
                      uint32_t tile_lo, uint32_t ntiles, const uint32_t* chunk_seg, uint32_t chunk_lo,
                      uint32_t nchunks, const void* in_keys, const uint32_t* in_vals, int pass,
                      uint32_t* hist, uint32_t* part, uint32_t* seg_valid, uint16_t* out_keys,
                      uint32_t* out_vals, hipStream_t st, bool count = true);
void launch_fill_zero(void* p, uint64_t bytes, hipStream_t st);
// lanes of k_accumulate resident at once on the device (CUs x workgroups per
// CU at its register use x 256): a lane count just above it leaves a second,
// nearly empty round of waves
uint64_t accumulate_capacity(int device, CurveId cv = CurveId::Bn254);
// after a part's sort: its entries per lane (*T) and the first lane of each
// of its segments (seg_lane0[seg_lo .. seg_lo + nseg)) from the segments'
// valid-entry counts, within the part's lanes [lane_lo, lane_hi)
void launch_lane_plan(const SegDesc* segs, const uint32_t* seg_valid, uint32_t seg_lo, uint32_t nseg,
                      uint64_t lane_lo, uint64_t lane_hi, uint64_t* seg_lane0, uint32_t* T, hipStream_t st);
// lanes [lane_lo, lane_hi) of the call's lane space: the part of segments
// [seg_lo, seg_lo + nseg), laid out by launch_lane_plan
void launch_accumulate(CurveId cv, const SegDesc* segs, const uint64_t* seg_lane0, uint32_t seg_lo,
                       uint32_t nseg, const MsmDesc* msms, const uint32_t* seg_valid, const uint16_t* keys,
                       const uint32_t* vals, const void* table, const uint32_t* T, uint64_t lane_lo,
                       uint64_t lane_hi, void* buckets, uint32_t* rec_keys, void* rec_pts, hipStream_t st);
// counts[0] / counts[1]: long / mid run counts (zeroed by the caller);
// lists: 2 * list_cap words (long list, then mid list)
void launch_merge_runs(CurveId cv, const uint32_t* keys, const void* pts, uint64_t nrec, uint32_t limit,
                       void* dest, uint32_t* counts, uint32_t* lists, uint32_t list_cap,
                       hipStream_t st);
// bits: 16 points per segment (U and the bit-plane sums), in the host form
// (Xyzz, 2^256 Montgomery, 128 B) for the host tail or in the stored form
// (144 B) for launch_tail
// segments [seg_lo, seg_hi) and their row/column waves [wave_lo, wave_hi)
void launch_bucket_reduce(CurveId cv, const SegDesc* segs, const uint64_t* seg_wave0,
                          const uint64_t* seg_rc0, uint32_t nseg, uint32_t seg_lo, uint32_t seg_hi,
                          uint64_t wave_lo, uint64_t wave_hi, uint32_t per, const void* buckets, void* rc,
                          void* bits, bool host_form, hipStream_t st);
// window combine and normalisation on the device (BN254 G1 and Grumpkin):
// bits in the stored form, wins: nseg stored points (scratch), out: nmsm
// results of 96 B (normalised Jacobian X | Y | Z, 2^256 Montgomery form)
void launch_tail(CurveId cv, const SegDesc* segs, const MsmDesc* msms, uint32_t nseg, uint32_t nmsm,
                 const void* bits, void* wins, void* out, hipStream_t st);
void launch_gen_bases(const void* ks, const void* gtab, uint64_t n, void* table,
                      hipStream_t st);
// out[i] = k_i * G on Grumpkin, 64-byte affine points (x, y in Fr Montgomery form)
// out[i] = k_i * G2, 128-byte affine points (x.c0, x.c1, y.c0, y.c1, Fp Montgomery)
void launch_gen_points_g2(const void* ks, const void* gtab, uint64_t n, void* out, hipStream_t st);
// BN254 G2 primitives on 256-byte host-form XYZZ points (see k_test_g2)
void launch_test_g2(int op, const void* p, const void* q, void* out, uint64_t n, hipStream_t st);
void launch_gen_points_grumpkin(const void* ks, const void* gtab, uint64_t n, void* out,
                                hipStream_t st);
void launch_test_fp(int op, const void* a, const void* b, void* out, uint64_t n,
                    hipStream_t st);
void launch_test_xyzz(int op, const void* p, const void* q, void* out, uint64_t n,
                      hipStream_t st);
void launch_test_glv(const void* scalars, void* out, uint64_t n, hipStream_t st);
void launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t st);
void launch_probe(int mode, uint32_t blocks, uint32_t iters, uint64_t* out, hipStream_t st);

}  // namespace zk
