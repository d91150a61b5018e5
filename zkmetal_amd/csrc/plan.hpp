// Work descriptors shared by the host planner and the kernels.
//
// A launch processes a batch of MSMs.  Each MSM j has n_j scalars, window
// width c_j and W_j = ceil(128 / c_j) signed windows; its GLV half-scalars
// address a base table laid out as [P_0, phi(P_0), P_1, phi(P_1), ...]
// (affine, 128 bytes per entry), so "entry" 2i+h of an MSM is base i
// (h = 0) or its endomorphism image (h = 1).  A segment is a set of digit
// entries that are sorted by bucket, accumulated into 2^(c_j - 1) buckets and
// reduced to one window sum.  Plain plan: every (MSM, window) pair is a
// segment of 2 n_j entries, and the window sums combine as sum_w 2^(c w) S_w.
// Window-shifted plan (a resident table that also holds the copies
// 2^(c w) P_i and their phi images for every window w, at entry w * stride):
// a segment holds the digits of wps consecutive windows, each window's 2 n_j
// entries padded to whole sort tiles, and its digits address the shifted
// copies, so the segments' sums simply add up.
#pragma once
#include <stdint.h>

namespace zk {

// The curves the engine runs MSMs on: BN254 G1 (coordinates in Fp, scalars
// mod r, GLV: two digit entries per scalar over the bases P and phi(P), two
// 128-bit half-scalars), Grumpkin (y^2 = x^3 - 17 over Fr, scalars mod p, no
// GLV: one entry per scalar, 255-bit windows) and BN254 G2 (the twist over
// Fp2, scalars mod r, no GLV, coordinates twice as wide).
enum class CurveId : uint32_t { Bn254 = 0, Grumpkin = 1, Bn254G2 = 2 };
// digit entries per scalar
constexpr uint32_t entries_per_point(CurveId cv) { return cv == CurveId::Bn254 ? 2u : 1u; }
constexpr uint32_t scalar_window_span(CurveId cv) { return cv == CurveId::Bn254 ? 128u : 255u; }
// windows per MSM the decomposition's histogram counters hold (c >= 2 for
// BN254 G1, c >= 4 for the 255-bit windows)
constexpr uint32_t kMaxWindows = 64;

// bytes per base-table slot (one affine point in the kernels' field form, or
// the c0 / c1 half of a G2 point)
constexpr uint64_t kTableEntryBytes = 128;
// bytes per stored XYZZ point over one field (buckets, merge records,
// reduction partials); a G2 point is two of them
constexpr uint64_t kBucketBytes = 144;

// table slots per digit entry's base, and per point (P and phi(P) for G1)
constexpr uint32_t slots_per_entry(CurveId cv) { return cv == CurveId::Bn254G2 ? 2u : 1u; }
constexpr uint32_t table_slots_per_point(CurveId cv) { return cv == CurveId::Grumpkin ? 1u : 2u; }
// stored point, host-form point (XYZZ in the 2^256 form) and result
// (Jacobian X | Y | Z) sizes
constexpr uint64_t point_bytes(CurveId cv) { return cv == CurveId::Bn254G2 ? 2 * kBucketBytes : kBucketBytes; }
constexpr uint64_t host_point_bytes(CurveId cv) { return cv == CurveId::Bn254G2 ? 256u : 128u; }
constexpr uint64_t result_bytes(CurveId cv) { return cv == CurveId::Bn254G2 ? 192u : 96u; }
// affine input point: x, y (64 B), or x.c0, x.c1, y.c0, y.c1 for G2 (128 B)
constexpr uint64_t input_point_bytes(CurveId cv) { return cv == CurveId::Bn254G2 ? 128u : 64u; }

struct MsmDesc {
    uint64_t scalar0;  // first scalar of this MSM in the batch scalar buffer
    uint64_t table0;   // first table entry (affine point index)
    uint32_t n;        // number of bases / scalars
    uint32_t c;        // window width in bits (<= 16)
    uint32_t nwin;     // ceil(128 / c) digit windows
    uint32_t seg0;     // first segment
    uint32_t form;     // scalar encoding: 0 standard form, 1 Fr Montgomery form
    uint32_t nseg;     // segments (nwin plain; ceil(nwin / wps) window-shifted)
    uint32_t wps;      // windows per segment (1 plain)
    uint32_t cstep;    // doublings between consecutive segments' sums (c plain, 0 shifted)
};

struct SegDesc {
    uint64_t ent0;    // first entry in the digit / sorted arrays
    uint64_t bkt0;    // first bucket in the bucket array
    uint64_t lane0;   // first accumulation lane
    uint64_t grp0;    // first bucket-reduction group
    uint32_t len;     // entries: 2 n plain, wps windows of padded 2 n shifted
    uint32_t nbk;     // buckets = 2^(c-1)
    uint32_t tile0;   // first sort tile
    uint32_t ntiles;  // sort tiles of this segment
    uint32_t msm;     // owning MSM
    uint32_t wadj;    // window-shifted: table stride - padded window length (0 plain)
    uint32_t chunk0;  // first histogram-scan chunk (kScanChunk counts each)
    uint32_t nchunks;
    uint32_t wbase;   // window-shifted: table entry of the segment's first window (0 plain)
    uint32_t pad;
};
// A digit at position pos of segment s, written for the segment's local
// window v (digit word bits 16..23), addresses table entry
// pos + v * wadj + wbase relative to the MSM's table0: for a plain segment
// (v = 0, wadj = wbase = 0) the position itself.

// Merge records: key bit 31 set = no payload (padding); bits 0..30 = index of
// the bucket (or window sum) the payload belongs to.
#define ZK_REC_INVALID 0x80000000u
#define ZK_REC_KEYMASK 0x7fffffffu

// sort tile: 4096 entries, 512 threads of 8 (measured against 8192 and 2048
// entries per tile: the scatter's 34 KB of LDS lets 4 tiles share a CU)
constexpr int kSortThreads = 512;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanChunk = 4096;  // histogram counts per scan block (16 per thread)

}  // namespace zk
