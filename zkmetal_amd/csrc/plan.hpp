// Work descriptors shared by the host planner and the kernels.
//
// A launch processes a batch of MSMs.  Each MSM j has n_j scalars, window
// width c_j and W_j = ceil(128 / c_j) signed windows; its GLV half-scalars
// address a base table laid out as [P_0, phi(P_0), P_1, phi(P_1), ...]
// (affine, 128 bytes per entry), so "entry" 2i+h of an MSM is base i
// (h = 0) or its endomorphism image (h = 1).  Every (MSM, window) pair is a
// segment: 2 n_j digit entries that are sorted by bucket, accumulated into
// 2^(c_j - 1) buckets and reduced to one window sum.
#pragma once
#include <stdint.h>

namespace zk {

// The two curves of the BN254 cycle the engine runs MSMs on: BN254 G1
// (coordinates in Fp, scalars mod r, GLV: two table entries per base, P and
// phi(P), and two 128-bit half-scalars) and Grumpkin (y^2 = x^3 - 17 over Fr,
// scalars mod p, no GLV: one entry per base, 255-bit scalars).
enum class CurveId : uint32_t { Bn254 = 0, Grumpkin = 1 };
constexpr uint32_t entries_per_point(CurveId cv) { return cv == CurveId::Bn254 ? 2u : 1u; }
constexpr uint32_t scalar_window_span(CurveId cv) { return cv == CurveId::Bn254 ? 128u : 255u; }
// windows per MSM the decomposition's histogram counters hold (c >= 2 for
// BN254 G1, c >= 4 for Grumpkin)
constexpr uint32_t kMaxWindows = 64;

// bytes per base-table entry (one affine point in the kernels' field form)
constexpr uint64_t kTableEntryBytes = 128;
// bytes per stored XYZZ point (buckets, merge records, reduction partials)
constexpr uint64_t kBucketBytes = 144;

struct MsmDesc {
    uint64_t scalar0;  // first scalar of this MSM in the batch scalar buffer
    uint64_t table0;   // first table entry (affine point index)
    uint32_t n;        // number of bases / scalars
    uint32_t c;        // window width in bits (<= 16)
    uint32_t nwin;     // ceil(128 / c)
    uint32_t seg0;     // first segment
    uint32_t form;     // scalar encoding: 0 standard form, 1 Fr Montgomery form
    uint32_t reserved;
};

struct SegDesc {
    uint64_t ent0;    // first entry in the digit / sorted arrays
    uint64_t bkt0;    // first bucket in the bucket array
    uint64_t lane0;   // first accumulation lane
    uint64_t grp0;    // first bucket-reduction group
    uint32_t len;     // entries = 2 n
    uint32_t nbk;     // buckets = 2^(c-1)
    uint32_t tile0;   // first sort tile
    uint32_t ntiles;  // sort tiles of this segment
    uint32_t msm;     // owning MSM
    uint32_t win;     // window index
    uint32_t chunk0;  // first histogram-scan chunk (kScanChunk counts each)
    uint32_t nchunks;
};

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
