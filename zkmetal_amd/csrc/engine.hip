// MSM engine: planner, launch sequence, host tail.
//
// Reference counterpart: MetalMSM.msm (Sources/zkMetal/MSM/MSMEngine.swift:287-909)
// and the batch loop _msm_batch_impl (Sources/zkMetal-ffi/zkMetal_ffi.swift:1241-1298).
// Differences by design:
//   * every n runs on the GPU (no n <= 2048 CPU route, MSMEngine.swift:295-297);
//   * a batch of MSMs is one launch sequence over all (MSM, window) segments
//     instead of a sequential loop;
//   * bases can stay resident on the device (the reference re-copies them on
//     every call, MSMEngine.swift:395-397).
#include "engine.hpp"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>

#include "gen.hpp"
#include "kernels.hpp"
#include "tuned_gfx950.hpp"

namespace zk {

#define ZK_TRY(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return kGpuError;       \
    } while (0)
#define ZK_ALLOC(buf, bytes)                          \
    do {                                              \
        if ((buf).ensure(bytes) != hipSuccess) return kAllocFailed; \
    } while (0)

hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    size_t want = std::max<size_t>(bytes, 256);
    if (acct && acct->limit && acct->used - cap + want > acct->limit) return hipErrorOutOfMemory;
    release();
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = want;
    if (acct) acct->used += want;
    return hipSuccess;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    if (acct) acct->used -= cap;
    p = nullptr;
    cap = 0;
}

hipError_t HostBuf::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    release();
    size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = want;
    return hipSuccess;
}

void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

// ------------------------------------------------------------------ planner

uint32_t default_window_bits(uint64_t n, CurveId cv) {
    // field multiplications: W * (entries * 10 madd + 2^(c-1) * ~40 for bucket
    // reduction), W = ceil(span / c) windows of entries_per_point * n entries
    uint32_t best = 16;
    double best_cost = 1e300;
    for (uint32_t c = 4; c <= 16; c++) {
        double W = (scalar_window_span(cv) + c - 1) / c;
        double cost = W * (10.0 * entries_per_point(cv) * (double)n + 40.0 * (double)(1u << (c - 1)));
        if (cost < best_cost) {
            best_cost = cost;
            best = c;
        }
    }
    return best;
}

// empty plan that keeps its arrays' capacity (the no-allocation path plans
// into arrays reserved at engine creation)
static void plan_reset(Plan& pl) {
    pl.msms.clear();
    pl.msm_scalar0.clear();
    pl.msm_dtile0.clear();
    pl.segs.clear();
    pl.seg_lane0.clear();
    pl.seg_grp0.clear();
    pl.seg_rcw0.clear();
    pl.tile_seg.clear();
    pl.chunk_seg.clear();
    pl.scalars = pl.entries = pl.buckets = pl.lanes = pl.groups = pl.rc_waves = 0;
    pl.tiles = pl.chunks = pl.T = pl.dtiles = 0;
    pl.nparts = 1;
    pl.curve = CurveId::Bn254;
    pl.G = 8;
}

// parts for a call of `entries` sorted entries: the exposed sort is the first
// part's, and every part adds a round of launches and merge records
uint32_t default_parts(uint64_t entries) {
    // measured: a 2^22-point MSM (2^26 entries) runs best in 2 parts, the
    // 64-MSM batch (2^27.7 entries) in 3; more parts add merge records and
    // launches faster than they hide sort time
    return entries >= (1ull << 27) ? 3u : entries >= (1ull << 24) ? 2u : 1u;
}

// Window-shifted or plain, by an estimate in mixed-addition units: every
// digit entry costs one, every bucket of the reduction kReduceCost (the
// row/column and bit-plane kernels spend about 3.3 times a bucket
// accumulation's time per bucket in large calls).  Shifted: ceil(128 / c_s)
// windows of 2n entries padded to whole sort tiles into one set of
// 2^(c_s - 1) buckets; plain: ceil(128 / c) windows of 2n entries, each with
// 2^(c-1) buckets.
bool shifted_pays(uint64_t n, uint32_t c, uint32_t c_s) {
    constexpr double kReduceCost = 3.3;
    const double padded = (double)((2 * n + kSortTile - 1) / kSortTile * kSortTile);
    const double plain = (double)((128 + c - 1) / c) * (2.0 * (double)n + kReduceCost * (double)(1u << (c - 1)));
    const double shifted = (double)((128 + c_s - 1) / c_s) * padded + kReduceCost * (double)(1u << (c_s - 1));
    return shifted < plain;
}

// A call of one MSM has a latency-bound reduction (about 0.16 ms from 2^16
// to 2^22 points whatever its bucket count), and its shifted buckets are
// deep enough to span several accumulation lanes, which costs merge work:
// measured on MI355X, one shifted MSM takes 0.42 / 0.74 / 1.81 / 6.23 ms
// against 0.61 / 0.65 / 1.71 / 6.05 ms plain at 2^16 / 2^18 / 2^20 / 2^22.
constexpr uint64_t kShiftSingleMax = 1u << 17;

// ceil(log2 x), at most 32
static uint32_t size_class(uint64_t x) {
    uint32_t k = 0;
    while (k < 32 && (1ull << k) < x) k++;
    return k;
}

int make_plan(const uint32_t* ns, const uint64_t* table0, uint32_t nmsm, const uint32_t* cs,
              Plan& pl, uint32_t form, uint64_t lane_cap, CurveId cv, uint32_t nparts,
              uint64_t shift_stride, uint32_t shift_c, uint32_t shift_floor, const Schedule* sched) {
    // parts for a call of `entries` entries: the request, else the
    // schedule's entry for its size class, else by size
    auto parts_for = [&](uint64_t entries) -> uint32_t {
        if (nparts) return nparts;
        if (sched && sched->parts[size_class(entries)]) return sched->parts[size_class(entries)];
        return default_parts(entries);
    };
    plan_reset(pl);
    pl.curve = cv;
    const uint32_t eps = entries_per_point(cv);
    const bool can_shift = cv == CurveId::Bn254 && shift_stride && shift_c >= 2 && shift_c <= 16 &&
                           shift_stride % kSortTile == 0;
    // a call of fewer MSMs than pipeline parts splits each window-shifted MSM's
    // windows over several segments, so that the parts have segments to take
    uint32_t spm = 1;
    if (can_shift && nmsm) {
        uint64_t est = 0;
        for (uint32_t j = 0; j < nmsm; j++) est += (uint64_t)eps * ns[j] * ((128 + shift_c - 1) / shift_c);
        const uint32_t P = parts_for(est);
        if (nmsm < P) spm = (P + nmsm - 1) / nmsm;
    }
    pl.msms.resize(nmsm);
    pl.msm_scalar0.resize(nmsm + 1);
    pl.msm_dtile0.resize(nmsm + 1);
    for (uint32_t j = 0; j < nmsm; j++) {
        if (ns[j] == 0) return kInvalidInput;
        uint32_t c = (cs && cs[j]) ? cs[j] : default_window_bits(ns[j], cv);
        if (c < 2 || c > 16) return kInvalidInput;
        const bool sh = can_shift && ((shift_floor && ns[j] >= shift_floor) ||
                                      (shifted_pays(ns[j], c, shift_c) && (nmsm > 1 || ns[j] <= kShiftSingleMax)));
        if (sh) c = shift_c;
        MsmDesc& m = pl.msms[j];
        m.scalar0 = pl.scalars;
        m.table0 = table0[j];
        m.n = ns[j];
        m.c = c;
        m.nwin = (scalar_window_span(cv) + c - 1) / c;
        if (m.nwin > kMaxWindows) return kInvalidInput;  // Grumpkin: c >= 4
        m.seg0 = (uint32_t)pl.segs.size();
        m.form = form;
        m.wps = sh ? (m.nwin + spm - 1) / spm : 1;
        m.nseg = (m.nwin + m.wps - 1) / m.wps;
        m.cstep = sh ? 0 : c;
        const uint32_t ntw = (eps * ns[j] + kSortTile - 1) / kSortTile;  // sort tiles per window
        pl.msm_scalar0[j] = pl.scalars;
        pl.msm_dtile0[j] = pl.dtiles;
        pl.dtiles += ntw;
        pl.scalars += ns[j];
        for (uint32_t g = 0; g < m.nseg; g++) {
            const uint32_t w0 = g * m.wps, nv = std::min(m.wps, m.nwin - w0);
            SegDesc s{};
            s.ent0 = pl.entries;
            s.bkt0 = pl.buckets;
            s.len = sh ? nv * ntw * (uint32_t)kSortTile : eps * ns[j];
            s.nbk = 1u << (c - 1);
            s.tile0 = pl.tiles;
            s.ntiles = sh ? nv * ntw : ntw;
            s.msm = j;
            s.wadj = sh ? (uint32_t)(shift_stride - (uint64_t)ntw * kSortTile) : 0u;
            s.wbase = sh ? (uint32_t)(w0 * shift_stride) : 0u;
            s.chunk0 = pl.chunks;
            s.nchunks = (256u * s.ntiles + kScanChunk - 1) / kScanChunk;
            pl.chunks += s.nchunks;
            pl.entries += s.len;
            pl.buckets += s.nbk;
            pl.tiles += s.ntiles;
            pl.segs.push_back(s);
        }
    }
    pl.msm_scalar0[nmsm] = pl.scalars;
    pl.msm_dtile0[nmsm] = pl.dtiles;
    // 32-bit positions / bucket ids on the device
    if (pl.entries >= (1ull << 31) || pl.buckets >= (1ull << 30)) return kInvalidInput;
    const uint32_t nseg = (uint32_t)pl.segs.size();
    // parts: consecutive segment ranges of about entries / P entries each
    uint32_t P = parts_for(pl.entries);
    P = std::max<uint32_t>(1, std::min<uint32_t>({P, kMaxParts, nseg}));
    {
        uint32_t s = 0;
        uint64_t done = 0;
        pl.nparts = 0;
        for (uint32_t k = 0; k < P && s < nseg; k++) {
            PlanPart& pt = pl.parts[pl.nparts++];
            pt = PlanPart{};
            pt.seg_lo = s;
            const uint64_t target = pl.entries * (k + 1) / P;
            // at least one segment, and leave one for each later part
            do {
                done += pl.segs[s].len;
                s++;
            } while (s < nseg - (P - 1 - k) && done < target);
            if (k == P - 1) {
                for (; s < nseg; s++) done += pl.segs[s].len;
            }
            pt.seg_hi = s;
        }
    }
    // accumulation, per part: at most lane_cap lanes (one resident round of
    // waves -- every segment rounds its lane count up, hence the - nseg), >= 4
    // entries each, and (up to 32) at least half a bucket's mean entry count,
    // so that a bucket's partials span few lanes and merge without a long-run
    // pass
    for (uint32_t k = 0; k < pl.nparts; k++) {
        PlanPart& pt = pl.parts[k];
        uint64_t ent = 0, bk = 0;
        for (uint32_t s = pt.seg_lo; s < pt.seg_hi; s++) {
            ent += pl.segs[s].len;
            bk += pl.segs[s].nbk;
        }
        const uint64_t ns_p = pt.seg_hi - pt.seg_lo;
        const uint64_t half_bucket = ent / (2 * std::max<uint64_t>(1, bk));
        const uint64_t cap = lane_cap > 2ull * ns_p + 64 ? lane_cap - ns_p : std::max<uint64_t>(lane_cap / 2, 64);
        pt.T = (uint32_t)std::max<uint64_t>(std::max<uint64_t>(4, std::min<uint64_t>(32, half_bucket)),
                                            (ent + cap - 1) / cap);
        pt.entries = ent;
        pt.tile_lo = pl.segs[pt.seg_lo].tile0;
        pt.tile_hi = pl.segs[pt.seg_hi - 1].tile0 + pl.segs[pt.seg_hi - 1].ntiles;
        pt.chunk_lo = pl.segs[pt.seg_lo].chunk0;
        pt.chunk_hi = pl.segs[pt.seg_hi - 1].chunk0 + pl.segs[pt.seg_hi - 1].nchunks;
    }
    pl.T = pl.parts[0].T;
    // bucket reduction: per buckets per lane before the shuffle tree; 8 keeps
    // the tree shallow for small calls, larger batches (throughput-bound) use
    // up to 32 so that fewer additions go to tree levels
    pl.G = 8;
    while (pl.G < 32 && (pl.buckets >> 17) >= 2 * pl.G) pl.G *= 2;
    if (sched && sched->per[size_class(pl.entries)]) pl.G = sched->per[size_class(pl.entries)];
    pl.seg_lane0.resize(nseg + 1);
    pl.seg_grp0.resize(nseg + 1);
    pl.seg_rcw0.resize(nseg + 1);
    uint32_t part = 0;
    for (uint32_t s = 0; s < nseg; s++) {
        SegDesc& sg = pl.segs[s];
        if (s == pl.parts[part].seg_hi) part++;
        PlanPart& pt = pl.parts[part];
        if (s == pt.seg_lo) {
            pt.lane_lo = pl.lanes;
            pt.rcw_lo = pl.rc_waves;
        }
        sg.lane0 = pl.lanes;
        sg.grp0 = pl.groups;
        pl.seg_lane0[s] = pl.lanes;
        pl.seg_grp0[s] = pl.groups;
        pl.seg_rcw0[s] = pl.rc_waves;
        pl.lanes += (sg.len + pt.T - 1) / pt.T;
        // bucket reduction: 2^h rows of 2^g buckets and 2^g columns of 2^h,
        // max(1, len/8) lanes per row / column, rows and columns wave-aligned
        const uint32_t cb = 31 - __builtin_clz(sg.nbk), g = cb / 2, h = cb - g;
        const uint64_t per = pl.G;
        const uint64_t pr = (1u << g) >= per ? (1u << g) / per : 1, pc = (1u << h) >= per ? (1u << h) / per : 1;
        pl.rc_waves += (((uint64_t)1 << h) * pr + 63) / 64 + (((uint64_t)1 << g) * pc + 63) / 64;
        pl.groups += (1u << h) + (1u << g);
        for (uint32_t t = 0; t < sg.ntiles; t++) pl.tile_seg.push_back(s);
        for (uint32_t t = 0; t < sg.nchunks; t++) pl.chunk_seg.push_back(s);
        if (s + 1 == pt.seg_hi) {
            pt.lane_hi = pl.lanes;
            pt.rcw_hi = pl.rc_waves;
        }
    }
    pl.seg_lane0[nseg] = pl.lanes;
    pl.seg_grp0[nseg] = pl.groups;
    pl.seg_rcw0[nseg] = pl.rc_waves;
    return kOk;
}

// ------------------------------------------------------------------ engine

int32_t Engine::create(Engine** out) {
    if (!out) return kInvalidInput;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return kNoGpu;
    }
    Engine* e = new Engine();
    if (hipGetDevice(&e->dev_) != hipSuccess || hipStreamCreateWithFlags(&e->stream_, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->stream_sort_, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return kGpuError;
    }
    {
        bool ok = hipEventCreate(&e->ev_dec_) == hipSuccess;
        for (uint32_t p = 0; p < kMaxParts; p++)
            for (hipEvent_t* ev : {&e->ev_s0_[p], &e->ev_s1_[p], &e->ev_a0_[p], &e->ev_a1_[p], &e->ev_m1_[p],
                                   &e->ev_r0_[p], &e->ev_r1_[p]})
                ok = ok && hipEventCreate(ev) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&e->dep_fork_, hipEventDisableTiming) == hipSuccess;
        for (uint32_t p = 0; p < kMaxParts; p++)
            ok = ok && hipEventCreateWithFlags(&e->dep_sort_[p], hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            delete e;
            return kGpuError;
        }
    }
    e->account_buffers();
    (void)e->noalloc_setup();  // on failure the no-allocation entry point reports kGpuError
    e->acc_cap_ = accumulate_capacity(e->dev_);
    e->acc_cap_g2_ = accumulate_capacity(e->dev_, CurveId::Bn254G2);
    memcpy(e->ctab_, kTunedPlain, sizeof(e->ctab_));
    memcpy(e->ctab_batch_, kTunedBatch, sizeof(e->ctab_batch_));
    for (int b = 0; b < 2; b++) {
        memcpy(e->sched_[b].parts, kTunedParts[b], sizeof(e->sched_[b].parts));
        memcpy(e->sched_[b].per, kTunedPer[b], sizeof(e->sched_[b].per));
    }
    for (auto& ev : e->ev_) {
        if (hipEventCreate(&ev) != hipSuccess) {
            delete e;
            return kGpuError;
        }
    }
    if (hipEventCreate(&e->ev_tail_) != hipSuccess) {
        delete e;
        return kGpuError;
    }
    for (auto& ev : e->stage_ev_) {
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            delete e;
            return kGpuError;
        }
    }
    *out = e;
    return kOk;
}

void Engine::account_buffers() {
    for (DevBuf* b : {&desc_, &digits_, &keysA_, &valsA_, &keysB_, &valsB_, &hist_, &part_, &segvalid_,
                      &buckets_, &recAk_, &recAp_, &recBk_, &rc_, &winsum_, &err_, &wins_, &res_,
                      &pts_in_, &ptable_, &scalars_, &srs_, &gtab_, &gtab_gk_, &gtab_g2_, &tmp_})
        b->acct = &ws_;
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (stream_sort_) (void)hipStreamSynchronize(stream_sort_);
    for (uint32_t p = 0; p < kMaxParts; p++)
        for (hipEvent_t ev : {ev_s0_[p], ev_s1_[p], ev_a0_[p], ev_a1_[p], ev_m1_[p], ev_r0_[p], ev_r1_[p],
                              dep_sort_[p]})
            if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : {ev_dec_, dep_fork_})
        if (ev) (void)hipEventDestroy(ev);
    if (stream_sort_) (void)hipStreamDestroy(stream_sort_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    for (auto& ev : ev_)
        if (ev) (void)hipEventDestroy(ev);
    if (ev_tail_) (void)hipEventDestroy(ev_tail_);
    for (auto& ev : stage_ev_)
        if (ev) (void)hipEventDestroy(ev);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// pack the plan's descriptor arrays into one blob (16-byte aligned sections)
struct DescLayout {
    size_t msms, msm_scalar0, msm_dtile0, segs, seg_lane0, seg_grp0, seg_rcw0, tile_seg, chunk_seg, part_T,
        total;
};

static DescLayout layout_for(const Plan& pl) {
    DescLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 15) & ~size_t(15);
        return o;
    };
    L.msms = take(pl.msms.size() * sizeof(MsmDesc));
    L.msm_scalar0 = take(pl.msm_scalar0.size() * 8);
    L.msm_dtile0 = take(pl.msm_dtile0.size() * 8);
    L.segs = take(pl.segs.size() * sizeof(SegDesc));
    L.seg_lane0 = take(pl.seg_lane0.size() * 8);
    L.seg_grp0 = take(pl.seg_grp0.size() * 8);
    L.seg_rcw0 = take(pl.seg_rcw0.size() * 8);
    L.tile_seg = take(pl.tile_seg.size() * 4);
    L.chunk_seg = take(pl.chunk_seg.size() * 4);
    L.part_T = take(kMaxParts * 4);  // entries per lane of each part (k_lane_plan)
    L.total = off;
    return L;
}

static constexpr uint32_t kRunLimit = 8;

// Host tail of a call: per MSM, result = sum_w 2^(c w) (U_w + sum_{t < c-1}
// 2^t Q_{w,t}) by one Horner over the bit positions b = c w + t (Q_{w,t} =
// Pl_t for t < g, else Ph_{t-g}; U_w at t = 0): ~c W doublings and W c
// additions per MSM; then the affine normalisation of all results with one
// inversion.  hw: the window planes as copied back; H: the curve's field.
// Host views of the window bit-plane sums: HostG1<H> for XYZZ over one field
// (128-byte host points, 96-byte results), HostG2View for BN254 G2 (256-byte
// host points: the c0 parts X, Y, ZZ, ZZZ, then the c1 parts; 192-byte
// results X | Y | Z, each c0 | c1).  Results are normalised: Z = 1, the
// identity is (1, 1, 0) (the reference's g2Identity for G2).
template <class H>
struct HostG1 {
    using Pt = zkh::Xyzz;
    static constexpr uint64_t kIn = 128, kOut = 96;
    static Pt load(const uint8_t* b) {
        Pt p;
        memcpy(&p, b, 128);
        return p;
    }
    static Pt identity() { return H::identity(); }
    static bool is_id(const Pt& p) { return H::is_id(p); }
    static Pt xdbl(const Pt& p) { return H::xdbl(p); }
    static Pt xadd(const Pt& a, const Pt& b) { return H::xadd(a, b); }
    static void finish(const std::vector<Pt>& res, uint8_t* out) {
        std::vector<zkh::Fp> xs, ys;
        H::batch_to_affine(res, xs, ys);
        const zkh::Fp one = H::one(), zero = zkh::zero();
        for (size_t j = 0; j < res.size(); j++) {
            uint8_t* o = out + j * kOut;
            const bool id = H::is_id(res[j]);
            memcpy(o, (id ? one : xs[j]).v, 32);
            memcpy(o + 32, (id ? one : ys[j]).v, 32);
            memcpy(o + 64, (id ? zero : one).v, 32);
        }
    }
};

struct HostG2View {
    using G = zkh::HostG2;
    using Pt = zkh::Xyzz2;
    static constexpr uint64_t kIn = 256, kOut = 192;
    static Pt load(const uint8_t* b) {
        Pt p;
        zkh::Fp* c0[4] = {&p.X.c0, &p.Y.c0, &p.ZZ.c0, &p.ZZZ.c0};
        zkh::Fp* c1[4] = {&p.X.c1, &p.Y.c1, &p.ZZ.c1, &p.ZZZ.c1};
        for (int k = 0; k < 4; k++) {
            memcpy(c0[k]->v, b + 32 * k, 32);
            memcpy(c1[k]->v, b + 128 + 32 * k, 32);
        }
        return p;
    }
    static Pt identity() { return G::identity(); }
    static bool is_id(const Pt& p) { return G::is_id(p); }
    static Pt xdbl(const Pt& p) { return G::xdbl(p); }
    static Pt xadd(const Pt& a, const Pt& b) { return G::xadd(a, b); }
    static void finish(const std::vector<Pt>& res, uint8_t* out) {
        const zkh::Fp2 one = G::one2(), zero = G::zero2();
        for (size_t j = 0; j < res.size(); j++) {
            uint8_t* o = out + j * kOut;
            zkh::Fp2 x, y, z = one;
            if (G::is_id(res[j])) {
                x = one;
                y = one;
                z = zero;
            } else {
                G::to_affine(res[j], x, y);
            }
            const zkh::Fp2* c[3] = {&x, &y, &z};
            for (int k = 0; k < 3; k++) {
                memcpy(o + 64 * k, c[k]->c0.v, 32);
                memcpy(o + 64 * k + 32, c[k]->c1.v, 32);
            }
        }
    }
};

template <class HC>
static void host_tail(const Plan& pl, const uint8_t* hw, uint8_t* out) {
    const uint32_t nmsm = (uint32_t)pl.msms.size();
    std::vector<typename HC::Pt> res(nmsm);
    auto tail = [&](uint32_t jb, uint32_t je) {
        for (uint32_t j = jb; j < je; j++) {
            const MsmDesc& m = pl.msms[j];
            const int cb = (int)m.c - 1;  // planes per segment (bucket index bits)
            typename HC::Pt acc = HC::identity();
            // Horner over the bits of sum_w 2^(cstep w) S_w: at bit b, segment
            // w adds its plane t = b - cstep w (and U_w at t = 0); plain
            // windows (cstep = c) meet one segment per bit, window-shifted
            // segments (cstep = 0) all of them
            for (int b = (int)(m.cstep * (m.nseg - 1)) + cb - 1; b >= 0; b--) {
                if (!HC::is_id(acc)) acc = HC::xdbl(acc);
                for (uint32_t w = 0; w < m.nseg; w++) {
                    const int t = b - (int)(m.cstep * w);
                    if (t < 0 || t >= cb) continue;
                    const uint8_t* bs = hw + (uint64_t)(m.seg0 + w) * 16 * HC::kIn;
                    acc = HC::xadd(acc, HC::load(bs + (uint64_t)(1 + t) * HC::kIn));  // Q_{w,t}
                    if (t == 0) acc = HC::xadd(acc, HC::load(bs));                 // U_w
                }
            }
            res[j] = acc;
        }
    };
    uint32_t nt = nmsm >= 16 ? std::min<uint32_t>(8, nmsm / 8) : 1;
    if (nt > 1) {
        std::vector<std::thread> th;
        uint32_t chunk = (nmsm + nt - 1) / nt;
        for (uint32_t t = 1; t < nt; t++)
            th.emplace_back(tail, std::min(nmsm, t * chunk), std::min(nmsm, (t + 1) * chunk));
        tail(0, std::min(nmsm, chunk));
        for (auto& x : th) x.join();
    } else {
        tail(0, nmsm);
    }
    HC::finish(res, out);
}

WsSizes ws_sizes(const Plan& pl, bool dev_tail) {
    WsSizes z{};
    const uint64_t nseg = pl.segs.size(), nmsm = pl.msms.size();
    const uint64_t ptb = point_bytes(pl.curve), hpb = host_point_bytes(pl.curve);
    const uint64_t rec_cap = 2 * pl.lanes + 2;
    z.desc = layout_for(pl).total;
    z.digits = pl.entries * 4;
    z.keys = pl.entries * 2;
    z.vals = pl.entries * 4;
    z.hist = (uint64_t)pl.tiles * 256 * 4;
    z.part = (uint64_t)pl.chunks * 4;
    z.segvalid = nseg * 4;
    z.buckets = pl.buckets * ptb;
    z.rec_keys = rec_cap * 4;
    z.rec_pts = rec_cap * ptb;
    z.rc = pl.groups * ptb;
    // bit-plane points: host form for the host tail, stored form for the device's
    z.bits = nseg * 16 * (dev_tail ? ptb : hpb);
    z.err = 16;  // the decomposition flag, the merge's long and mid run counters
    z.wins = dev_tail ? nseg * ptb : 0;
    z.res = dev_tail ? nmsm * result_bytes(pl.curve) : 0;
    // copied back: the results (device tail) or the bit planes, then err, then
    // the per-segment valid-entry counts
    z.back = dev_tail ? z.res : z.bits;
    z.host = z.back + z.err + nseg * 4;
    const uint64_t parts[] = {z.desc, z.digits, z.keys, z.vals, z.keys, z.vals, z.hist, z.part,
                              z.segvalid, z.buckets, z.rec_keys, z.rec_pts, z.rec_keys, z.rc, z.bits,
                              z.err, z.wins, z.res};
    for (uint64_t b : parts) z.device += (b + 255) & ~uint64_t(255);
    return z;
}

int32_t Engine::run(const Plan& pl, const void* d_scalars, const void* d_table, uint64_t table_n,
                    uint8_t* out96) {
    const uint32_t nmsm = (uint32_t)pl.msms.size();
    // window combine on the device (graph mode, large batches) or the host
    const bool dev_tail = device_tail(nmsm, pl.curve);
    const WsSizes z = ws_sizes(pl, dev_tail);
    ZK_ALLOC(desc_, z.desc);
    ZK_ALLOC(hdesc_, z.desc);
    ZK_ALLOC(digits_, z.digits);
    ZK_ALLOC(keysA_, z.keys);
    ZK_ALLOC(valsA_, z.vals);
    ZK_ALLOC(keysB_, z.keys);
    ZK_ALLOC(valsB_, z.vals);
    ZK_ALLOC(hist_, z.hist);
    ZK_ALLOC(part_, z.part);
    ZK_ALLOC(segvalid_, z.segvalid);
    ZK_ALLOC(buckets_, z.buckets);
    ZK_ALLOC(recAk_, z.rec_keys);
    ZK_ALLOC(recAp_, z.rec_pts);
    ZK_ALLOC(recBk_, z.rec_keys);
    ZK_ALLOC(rc_, z.rc);
    ZK_ALLOC(winsum_, z.bits);
    if (dev_tail) {
        ZK_ALLOC(wins_, z.wins);
        ZK_ALLOC(res_, z.res);
    }
    ZK_ALLOC(err_, z.err);
    ZK_ALLOC(hwin_, z.host);
    Workspace ws;
    ws.desc = (uint8_t*)desc_.p;
    ws.digits = digits_.p;
    ws.keysA = keysA_.p;
    ws.valsA = valsA_.p;
    ws.keysB = keysB_.p;
    ws.valsB = valsB_.p;
    ws.hist = hist_.p;
    ws.part = part_.p;
    ws.segvalid = segvalid_.p;
    ws.buckets = buckets_.p;
    ws.recAk = recAk_.p;
    ws.recAp = recAp_.p;
    ws.recBk = recBk_.p;
    ws.rc = rc_.p;
    ws.bits = winsum_.p;
    ws.err = err_.p;
    ws.wins = wins_.p;
    ws.res = res_.p;
    ws.hdesc = (uint8_t*)hdesc_.p;
    ws.hw = (uint8_t*)hwin_.p;
    return run_ws(pl, ws, z, dev_tail, graph_mode_, d_scalars, d_table, table_n, out96);
}

int32_t Engine::run_ws(const Plan& pl, const Workspace& ws, const WsSizes& z, bool dev_tail, int32_t graph_mode,
                       const void* d_scalars, const void* d_table, uint64_t table_n, uint8_t* out96) {
    auto t_wall0 = std::chrono::steady_clock::now();
    const uint32_t nmsm = (uint32_t)pl.msms.size();
    const uint32_t nseg = (uint32_t)pl.segs.size();
    const DescLayout L = layout_for(pl);
    const uint64_t ptb = point_bytes(pl.curve);
    const uint64_t back_bytes = z.back;

    uint8_t* hb = ws.hdesc;
    memcpy(hb + L.msms, pl.msms.data(), pl.msms.size() * sizeof(MsmDesc));
    memcpy(hb + L.msm_scalar0, pl.msm_scalar0.data(), pl.msm_scalar0.size() * 8);
    memcpy(hb + L.msm_dtile0, pl.msm_dtile0.data(), pl.msm_dtile0.size() * 8);
    memcpy(hb + L.segs, pl.segs.data(), pl.segs.size() * sizeof(SegDesc));
    memcpy(hb + L.seg_lane0, pl.seg_lane0.data(), pl.seg_lane0.size() * 8);
    memcpy(hb + L.seg_grp0, pl.seg_grp0.data(), pl.seg_grp0.size() * 8);
    memcpy(hb + L.seg_rcw0, pl.seg_rcw0.data(), pl.seg_rcw0.size() * 8);
    memcpy(hb + L.tile_seg, pl.tile_seg.data(), pl.tile_seg.size() * 4);
    memcpy(hb + L.chunk_seg, pl.chunk_seg.data(), pl.chunk_seg.size() * 4);
    for (uint32_t p = 0; p < kMaxParts; p++)
        reinterpret_cast<uint32_t*>(hb + L.part_T)[p] = p < pl.nparts ? pl.parts[p].T : 0u;
    uint8_t* db = ws.desc;
    const MsmDesc* d_msms = (const MsmDesc*)(db + L.msms);
    const uint64_t* d_msm_scalar0 = (const uint64_t*)(db + L.msm_scalar0);
    const uint64_t* d_msm_dtile0 = (const uint64_t*)(db + L.msm_dtile0);
    const SegDesc* d_segs = (const SegDesc*)(db + L.segs);
    // seg_lane0 and part_T: the host plan's values, rewritten on the device
    // per part by k_lane_plan from the sorted valid counts
    uint64_t* d_seg_lane0 = (uint64_t*)(db + L.seg_lane0);
    uint32_t* d_part_T = (uint32_t*)(db + L.part_T);
    const uint64_t* d_seg_grp0 = (const uint64_t*)(db + L.seg_grp0);
    const uint64_t* d_seg_rcw0 = (const uint64_t*)(db + L.seg_rcw0);
    const uint32_t* d_tile_seg = (const uint32_t*)(db + L.tile_seg);
    const uint32_t* d_chunk_seg = (const uint32_t*)(db + L.chunk_seg);
    uint32_t* d_err = (uint32_t*)ws.err;
    hipStream_t st = stream_;
    const uint32_t P = pl.nparts;
    // the merge lists (long and mid runs): a listed run has more than
    // kRunLimit records, so each list holds at most nrec / (kRunLimit + 1)
    // runs; both sit in recBk (2 * lanes + 2 words)
    const uint64_t nrec = 2 * pl.parts[P - 1].lane_hi;
    const uint32_t list_cap = (uint32_t)std::min<uint64_t>(nrec / (kRunLimit + 1) + 1, 1u << 29);
    uint8_t* hw = ws.hw;
    // The device part of the call.  Stream A (the engine stream): descriptors,
    // decomposition, then the accumulation of each part, one merge of every
    // part's records, the bucket reduction and the device tail; stream S: the
    // two sort passes and the lane plan of each part, so that part p + 1 sorts
    // while part p accumulates (the accumulation leaves registers and LDS for
    // the sort's workgroups).
    // Timing events are recorded only outside graph capture; the dependency
    // events are part of a captured graph.
    auto enqueue = [&](bool events) -> int32_t {
        auto rec = [&](hipEvent_t e, hipStream_t s) -> hipError_t { return events ? hipEventRecord(e, s) : hipSuccess; };
        // one part: everything on stream A (no cross-stream waits)
        hipStream_t S = P > 1 ? stream_sort_ : st;
        ZK_TRY(hipMemcpyAsync(ws.desc, ws.hdesc, L.total, hipMemcpyHostToDevice, st));
        launch_fill_zero(ws.err, z.err, st);
        ZK_TRY(rec(ev_[0], st));
        // the decomposition also counts the first sort pass's tile histograms
        launch_decompose(pl.curve, d_msms, d_msm_dtile0, nmsm, pl.dtiles, d_segs, d_scalars, d_table, table_n,
                         ws.digits, (uint32_t*)ws.hist, d_err, st);
        ZK_TRY(rec(ev_dec_, st));
        if (P > 1) {  // the sorts need the digits, not the cleared buckets
            ZK_TRY(hipEventRecord(dep_fork_, st));
            ZK_TRY(hipStreamWaitEvent(S, dep_fork_, 0));
        }
        launch_fill_zero(ws.buckets, pl.buckets * ptb, st);
        for (uint32_t p = 0; p < P; p++) {
            const PlanPart& pt = pl.parts[p];
            const uint32_t ns_p = pt.seg_hi - pt.seg_lo;
            ZK_TRY(rec(ev_s0_[p], S));
            launch_sort_pass(d_segs, pt.seg_lo, ns_p, d_tile_seg, pt.tile_lo, pt.tile_hi - pt.tile_lo, d_chunk_seg,
                             pt.chunk_lo, pt.chunk_hi - pt.chunk_lo, ws.digits, nullptr, 0, (uint32_t*)ws.hist,
                             (uint32_t*)ws.part, (uint32_t*)ws.segvalid, (uint16_t*)ws.keysA,
                             (uint32_t*)ws.valsA, S, false);
            launch_sort_pass(d_segs, pt.seg_lo, ns_p, d_tile_seg, pt.tile_lo, pt.tile_hi - pt.tile_lo, d_chunk_seg,
                             pt.chunk_lo, pt.chunk_hi - pt.chunk_lo, ws.keysA, (const uint32_t*)ws.valsA, 1,
                             (uint32_t*)ws.hist, (uint32_t*)ws.part, (uint32_t*)ws.segvalid,
                             (uint16_t*)ws.keysB, (uint32_t*)ws.valsB, S);
            launch_lane_plan(d_segs, (const uint32_t*)ws.segvalid, pt.seg_lo, ns_p, pt.lane_lo, pt.lane_hi,
                             d_seg_lane0, d_part_T + p, S);
            ZK_TRY(rec(ev_s1_[p], S));
            if (P > 1) ZK_TRY(hipEventRecord(dep_sort_[p], S));
        }
        for (uint32_t p = 0; p < P; p++) {
            const PlanPart& pt = pl.parts[p];
            if (P > 1) ZK_TRY(hipStreamWaitEvent(st, dep_sort_[p], 0));
            ZK_TRY(rec(ev_a0_[p], st));
            launch_accumulate(pl.curve, d_segs, d_seg_lane0, pt.seg_lo, pt.seg_hi - pt.seg_lo, d_msms,
                              (const uint32_t*)ws.segvalid, (const uint16_t*)ws.keysB, (const uint32_t*)ws.valsB,
                              d_table, d_part_T + p, pt.lane_lo, pt.lane_hi, ws.buckets, (uint32_t*)ws.recAk,
                              ws.recAp, st);
            ZK_TRY(rec(ev_a1_[p], st));
        }
        // bucket partials that crossed lane boundaries, every part's at once
        // after the last accumulation (a part's records and buckets are its
        // own, and records are sorted by bucket across parts): short runs by
        // their first lane, longer runs (heavy buckets) by lane groups.  One
        // merge after the last part instead of one after each keeps the
        // accumulations back to back.
        launch_merge_runs(pl.curve, (const uint32_t*)ws.recAk, (const uint8_t*)ws.recAp, nrec, kRunLimit,
                          ws.buckets, d_err + 1, (uint32_t*)ws.recBk, list_cap, st);
        ZK_TRY(rec(ev_m1_[0], st));
        // the bucket reduction of all parts at once, after the last merge: as
        // a kernel beside the next part's accumulation it finds no free slots
        // (its waves need as many registers as the accumulation's) and
        // measured slower
        ZK_TRY(rec(ev_r0_[0], st));
        launch_bucket_reduce(pl.curve, d_segs, d_seg_rcw0, d_seg_grp0, nseg, 0, nseg, 0, pl.rc_waves, pl.G,
                             ws.buckets, ws.rc, ws.bits, !dev_tail, st);
        ZK_TRY(rec(ev_r1_[0], st));
        if (dev_tail) {
            launch_tail(pl.curve, d_segs, d_msms, nseg, nmsm, ws.bits, ws.wins, ws.res, st);
            ZK_TRY(rec(ev_tail_, st));
        }
        ZK_TRY(rec(ev_[kStCopyOut], st));
        ZK_TRY(hipMemcpyAsync(hw, dev_tail ? ws.res : ws.bits, back_bytes, hipMemcpyDeviceToHost, st));
        // err[0]: decomposition error flag; err[1], err[2]: long / mid run
        // counts of the merge (above list_cap: listed runs were dropped)
        ZK_TRY(hipMemcpyAsync(hw + back_bytes, ws.err, 4 * 3, hipMemcpyDeviceToHost, st));
        ZK_TRY(hipMemcpyAsync(hw + back_bytes + z.err, ws.segvalid, (uint64_t)nseg * 4,
                              hipMemcpyDeviceToHost, st));
        return kOk;
    };
    // The shape key: a captured sequence is replayed when the call has the
    // same shape (the curve -- its kernels -- and every launch dimension the
    // capture baked in) and every buffer it touches is where it was at
    // capture.  Graph mode on: capture on a new key, else replay; auto: the
    // same from the second call in a row with one key (a one-off call runs
    // eagerly and pays no capture).
    bool graph = false;
    std::vector<uint64_t> key;
    if (graph_mode != kGraphOff) {
        key.reserve(pl.msms.size() * 3 + 40);
        key.push_back((uint64_t)pl.curve);
        for (const MsmDesc& m : pl.msms) {
            key.push_back(m.n);
            key.push_back(m.c);
            key.push_back(m.table0);
        }
        for (uint64_t v : {pl.scalars, pl.entries, pl.buckets, pl.lanes, pl.groups, pl.rc_waves,
                           (uint64_t)pl.tiles, (uint64_t)pl.chunks, (uint64_t)pl.T, (uint64_t)pl.dtiles,
                           (uint64_t)pl.G, (uint64_t)nseg})
            key.push_back(v);
        key.push_back(table_n);  // places the table's identity bitmap
        for (uint32_t p = 0; p < P; p++) {
            const PlanPart& pt = pl.parts[p];
            for (uint64_t v : {(uint64_t)pt.seg_hi, (uint64_t)pt.T, pt.lane_hi, pt.rcw_hi}) key.push_back(v);
        }
        key.push_back(dev_tail ? 1 : 0);
        for (const void* p : std::initializer_list<const void*>{
                 d_scalars, d_table, ws.desc, ws.hdesc, ws.digits, ws.keysA, ws.valsA, ws.keysB, ws.valsB,
                 ws.hist, ws.part, ws.segvalid, ws.buckets, ws.recAk, ws.recAp, ws.recBk, ws.rc, ws.bits,
                 ws.err, ws.hw, ws.wins, ws.res})
            key.push_back((uint64_t)(uintptr_t)p);
        graph = graph_mode == kGraphOn || key == last_key_ || (graph_exec_ && key == graph_key_);
        last_key_ = key;
    }
    if (graph) {
        if (!graph_exec_ || key != graph_key_) {
            if (graph_exec_) {
                (void)hipGraphExecDestroy(graph_exec_);
                graph_exec_ = nullptr;
            }
            hipGraph_t g = nullptr;
            ZK_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int32_t rc = enqueue(false);
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc != kOk || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                return kGpuError;
            }
            const hipError_t ei = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                graph_exec_ = nullptr;
                (void)hipGetLastError();
                return kGpuError;
            }
            graph_key_ = key;
            graph_captures_++;
        }
        ZK_TRY(hipEventRecord(ev_[0], st));
        ZK_TRY(hipGraphLaunch(graph_exec_, st));
        ZK_TRY(hipEventRecord(ev_[kNumStages], st));
    } else {
        const int32_t rc = enqueue(true);
        if (rc != kOk) return rc;
        ZK_TRY(hipEventRecord(ev_[kNumStages], st));
    }
    ZK_TRY(hipStreamSynchronize(st));
    ZK_TRY(hipGetLastError());
    uint32_t err[3];
    memcpy(err, hw + back_bytes, 4 * 3);
    if (err[0]) return kGpuError;
    if (err[1] > list_cap || err[2] > list_cap) return kGpuError;

    // stage times: each stage's kernels summed over the parts (they overlap
    // across streams; a graph replay has only the device total)
    for (int s = 0; s < kNumStages; s++) timing_.stage_ms[s] = 0;
    timing_.device_tail_ms = 0;
    if (!graph) {
        auto span = [](hipEvent_t a, hipEvent_t b) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            return ms;
        };
        timing_.stage_ms[kStDecompose] = span(ev_[0], ev_dec_);
        for (uint32_t p = 0; p < P; p++) {
            timing_.stage_ms[kStSort] += span(ev_s0_[p], ev_s1_[p]);
            timing_.stage_ms[kStAccumulate] += span(ev_a0_[p], ev_a1_[p]);
        }
        timing_.stage_ms[kStMerge] = span(ev_a1_[P - 1], ev_m1_[0]);
        timing_.stage_ms[kStReduce] = span(ev_r0_[0], ev_r1_[0]);
        if (dev_tail) timing_.device_tail_ms = span(ev_r1_[0], ev_tail_);
        timing_.stage_ms[kStCopyOut] = span(ev_[kStCopyOut], ev_[kNumStages]);
    }
    (void)hipEventElapsedTime(&timing_.device_ms, ev_[0], ev_[kNumStages]);

    // host tail: Horner over windows, then affine normalisation; the device
    // tail left the normalised results
    auto t_tail0 = std::chrono::steady_clock::now();
    if (dev_tail) memcpy(out96, hw, back_bytes);
    else if (pl.curve == CurveId::Grumpkin) host_tail<HostG1<zkh::HR>>(pl, hw, out96);
    else if (pl.curve == CurveId::Bn254G2) host_tail<HostG2View>(pl, hw, out96);
    else host_tail<HostG1<zkh::HP>>(pl, hw, out96);
    auto t_end = std::chrono::steady_clock::now();
    timing_.host_tail_ms = dev_tail ? 0.f : std::chrono::duration<float, std::milli>(t_end - t_tail0).count();
    timing_.wall_ms = std::chrono::duration<float, std::milli>(t_end - t_wall0).count();
    timing_.nmsm = nmsm;
    timing_.window_bits_first = nmsm ? pl.msms[0].c : 0;
    timing_.T = pl.T;
    timing_.G = pl.G;
    timing_.parts = pl.nparts;
    timing_.entries = pl.entries;
    timing_.buckets = pl.buckets;
    timing_.lanes = pl.lanes;
    uint64_t valid = 0;
    const uint32_t* sv = (const uint32_t*)(hw + back_bytes + z.err);
    for (uint32_t s = 0; s < nseg; s++) valid += sv[s];
    timing_.valid_entries = valid;
    return kOk;
}

// ------------------------------------------------------------ entry points

int32_t Engine::msm_host(const uint8_t* points, const uint8_t* scalars, uint32_t n,
                         uint8_t* out96) {
    uint32_t counts[1] = {n};
    return msm_batch_host(points, scalars, counts, 1, out96);
}

// Host-point calls (zkMetal_ffi.swift:1241-1298 takes points on every call):
// a batch whose point slices are each a prefix of the resident SRS runs on
// the resident table.  The check is exact: one 64-byte compare against the cached first
// SRS base rejects unrelated inputs at once; otherwise every slice is compared
// with the host image of the SRS (in parallel), which is far cheaper than
// sending the bases over PCIe and rebuilding their table.
bool Engine::srs_candidate(const uint8_t* points, const uint32_t* counts, uint32_t nmsm) {
    if (!srs_detect_ || srs_n_ == 0) return false;
    uint64_t off = 0, total = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        const uint64_t n = counts[j];
        if (n > srs_n_) return false;
        if (n == 0) continue;
        if (memcmp(points + off, srs_first_, 64) != 0) return false;
        off += n * 64;
        total += n;
    }
    if (total == 0) return false;
    if (srs_host_.size() != (uint64_t)srs_n_ * 64) {
        std::vector<uint8_t> img((uint64_t)srs_n_ * 64);
        if (srs_download(img.data(), srs_n_) != kOk) return false;
        srs_host_.swap(img);
    }
    return true;
}

bool Engine::srs_confirm(const uint8_t* points, const uint32_t* counts, uint32_t nmsm) {
    // split the bytes to compare into pieces for up to 16 host threads
    struct Piece { const uint8_t* a; const uint8_t* b; uint64_t bytes; };
    std::vector<Piece> pieces;
    uint64_t total = 0;
    for (uint32_t j = 0; j < nmsm; j++) total += (uint64_t)counts[j] * 64;
    const uint64_t step = std::max<uint64_t>(1ull << 22, total / 16 + 1);
    uint64_t off = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        const uint64_t bytes = (uint64_t)counts[j] * 64;
        for (uint64_t o = 0; o < bytes; o += step)
            pieces.push_back({points + off + o, srs_host_.data() + o, std::min(step, bytes - o)});
        off += bytes;
    }
    std::atomic<bool> same{true};
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i = next++; i < pieces.size() && same.load(std::memory_order_relaxed); i = next++)
            if (memcmp(pieces[i].a, pieces[i].b, pieces[i].bytes) != 0) same = false;
    };
    const uint32_t nt = (uint32_t)std::min<size_t>(16, pieces.size());
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
    return same.load();
}

int32_t Engine::stage_upload(void* dst, const uint8_t* src, uint64_t bytes) {
    if (bytes == 0) return kOk;
    const uint64_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
    const uint32_t T = (uint32_t)std::min<uint64_t>(kStageSlots, nchunks);
    for (uint32_t t = 0; t < T; t++)
        if (stage_[t].ensure(std::min<uint64_t>(kStageChunk, bytes)) != hipSuccess) return kAllocFailed;
    // thread t fills slot t with chunks t, t + T, ...: it reuses its slot once
    // the copy that last read it -- in this upload or an earlier one -- has
    // completed (the slot's event; one never recorded is complete)
    std::atomic<int32_t> err{kOk};
    auto worker = [&](uint32_t t) {
        for (uint64_t c = t; c < nchunks && err.load(std::memory_order_relaxed) == kOk; c += T) {
            if (hipEventSynchronize(stage_ev_[t]) != hipSuccess) {
                err = kGpuError;
                return;
            }
            const uint64_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
            memcpy(stage_[t].p, src + off, len);
            if (hipMemcpyAsync((uint8_t*)dst + off, stage_[t].p, len, hipMemcpyHostToDevice, stream_) != hipSuccess ||
                hipEventRecord(stage_ev_[t], stream_) != hipSuccess) {
                err = kGpuError;
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < T; t++) th.emplace_back(worker, t);
    worker(0);
    for (auto& x : th) x.join();
    return err.load();
}

int32_t Engine::srs_set_first() {
    srs_host_.clear();
    return srs_download(srs_first_, 1);
}

int32_t Engine::msm_batch_host(const uint8_t* points, const uint8_t* scalars,
                               const uint32_t* counts, uint32_t nmsm, uint8_t* out96, CurveId cv) {
    if (!points || !scalars || !counts || !out96 || nmsm == 0) return kInvalidInput;
    std::vector<uint32_t> ns;
    std::vector<uint64_t> t0;
    std::vector<uint32_t> idx;
    uint64_t total = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        if (counts[j] == 0) {
            memset(out96 + (uint64_t)j * result_bytes(cv), 0, result_bytes(cv));  // zkMetal_ffi.swift:1260-1264
            continue;
        }
        ns.push_back(counts[j]);
        t0.push_back(entries_per_point(cv) * total);
        idx.push_back(j);
        total += counts[j];
    }
    if (ns.empty()) return kOk;
    if (total >= (1ull << 30)) return kInvalidInput;
    // the scalars go up first, in pinned chunks; whichever base table the
    // call ends up using reads them from scalars_
    ZK_ALLOC(scalars_, total * 32);
    int32_t rc = stage_upload(scalars_.p, scalars, total * 32);
    if (rc != kOk) return rc;
    if (cv == CurveId::Bn254 && srs_candidate(points, counts, nmsm)) {
        // slices that start with the SRS's first base: run on the resident
        // table while host threads confirm the slices byte for byte; the
        // result stands only if they do
        bool same = false;
        std::thread cmp([&] { same = srs_confirm(points, counts, nmsm); });
        std::vector<uint8_t> res(nmsm * 96);
        rc = msm_table(srs_.p, srs_n_, scalars_.p, true, counts, nmsm, res.data());
        cmp.join();
        if (rc != kOk) return rc;
        if (same) {
            srs_hits_++;
            memcpy(out96, res.data(), res.size());
            return kOk;
        }
    }
    Plan pl;
    rc = make_plan(ns.data(), t0.data(), (uint32_t)ns.size(), window_bits_for(ns, cv).data(), pl, 0,
                   lane_cap(cv), cv, parts_req_, 0, 0, 0, schedule((uint32_t)ns.size()));
    if (rc != kOk) return rc;
    ZK_ALLOC(pts_in_, total * input_point_bytes(cv));
    ZK_ALLOC(ptable_, table_bytes(total, cv));
    rc = stage_upload(pts_in_.p, points, total * input_point_bytes(cv));
    if (rc != kOk) return rc;
    launch_build_table(cv, pts_in_.p, ptable_.p, total, stream_);
    std::vector<uint8_t> res(ns.size() * result_bytes(cv));
    rc = run(pl, scalars_.p, ptable_.p, total, res.data());
    if (rc != kOk) return rc;
    const uint64_t rb = result_bytes(cv);
    for (size_t k = 0; k < idx.size(); k++) memcpy(out96 + (uint64_t)idx[k] * rb, res.data() + k * rb, rb);
    return kOk;
}

// MultiMSM.swift:21-60 (k MSMs sharing one point set).  The reference runs
// the k MSMs one after another over the uploaded points; here the base table
// is built once and every scalar set is an MSM over its full prefix, all in
// one batched launch sequence.
int32_t Engine::multi_msm_host(const uint8_t* points, uint32_t n, const uint8_t* scalars,
                               uint32_t k, uint8_t* out96) {
    if (!points || !scalars || !out96 || n == 0 || k == 0) return kInvalidInput;
    if ((uint64_t)n * k >= (1ull << 30)) return kInvalidInput;
    ZK_ALLOC(pts_in_, (uint64_t)n * 64);
    ZK_ALLOC(ptable_, table_bytes(n));
    const int32_t rc = stage_upload(pts_in_.p, points, (uint64_t)n * 64);
    if (rc != kOk) return rc;
    launch_build_table(CurveId::Bn254, pts_in_.p, ptable_.p, n, stream_);
    std::vector<uint32_t> counts(k, n);
    return msm_table(ptable_.p, n, scalars, false, counts.data(), k, out96);
}

int32_t Engine::msm_device(const void* d_points, const void* d_scalars, const uint32_t* counts,
                           uint32_t nmsm, uint8_t* out96, CurveId cv) {
    if (!d_points || !d_scalars || !counts || !out96 || nmsm == 0) return kInvalidInput;
    std::vector<uint32_t> ns, idx;
    std::vector<uint64_t> t0;
    uint64_t total = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        if (counts[j] == 0) {
            memset(out96 + (uint64_t)j * result_bytes(cv), 0, result_bytes(cv));
            continue;
        }
        ns.push_back(counts[j]);
        t0.push_back(entries_per_point(cv) * total);
        idx.push_back(j);
        total += counts[j];
    }
    if (ns.empty()) return kOk;
    if (total >= (1ull << 30)) return kInvalidInput;
    // the table build goes first: the host plans the call while it runs
    ZK_ALLOC(ptable_, table_bytes(total, cv));
    launch_build_table(cv, d_points, ptable_.p, total, stream_);
    Plan& pl = plan_;
    int rc = make_plan(ns.data(), t0.data(), (uint32_t)ns.size(), window_bits_for(ns, cv).data(), pl, 0,
                       lane_cap(cv), cv, parts_req_, 0, 0, 0, schedule((uint32_t)ns.size()));
    if (rc != kOk) return rc;
    std::vector<uint8_t> res(ns.size() * result_bytes(cv));
    rc = run(pl, d_scalars, ptable_.p, total, res.data());
    if (rc != kOk) return rc;
    const uint64_t rb = result_bytes(cv);
    for (size_t k = 0; k < idx.size(); k++) memcpy(out96 + (uint64_t)idx[k] * rb, res.data() + k * rb, rb);
    return kOk;
}

int32_t Engine::srs_alloc(uint64_t n, uint64_t* stride) {
    *stride = 0;
    srs_shift_ = 0;
    const uint64_t base = table_bytes(n);
    if (shift_mode_ != kShiftOff) {
        // window copies kSortTile-aligned (a padded window's positions stay
        // inside its copy), all table entries below 2^31 (32-bit sort values)
        const uint64_t s = ((base + kTableEntryBytes - 1) / kTableEntryBytes + kSortTile - 1) / kSortTile * kSortTile;
        const uint64_t nw = (128 + kShiftC - 1) / kShiftC;
        if (s * nw < (1ull << 31) && srs_.ensure(s * nw * kTableEntryBytes) == hipSuccess) {
            *stride = s;
            return kOk;
        }
    }
    ZK_ALLOC(srs_, base);
    return kOk;
}

int32_t Engine::srs_finish(uint64_t n, uint64_t stride) {
    if (stride) launch_shift_windows(srs_.p, n, stride, kShiftC, (128 + kShiftC - 1) / kShiftC, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipGetLastError());
    srs_shift_ = stride;
    return kOk;
}

uint32_t Engine::shift_floor(uint32_t nmsm) const {
    if (shift_mode_ == kShiftAll) return 1;
    return device_tail(nmsm, CurveId::Bn254) ? kShiftDeviceTailFloor : 0;
}

uint64_t Engine::plan_shift(const void* table, uint64_t shift_stride) const {
    if (c_override_ || shift_mode_ == kShiftOff) return 0;  // a forced window width runs plain
    if (shift_stride == kOwnTable) return table == srs_.p ? srs_shift_ : 0;
    return shift_stride;
}

int32_t Engine::srs_upload(const uint8_t* points, uint32_t n) {
    if (!points || n == 0) return kInvalidInput;
    if (servers_) return kInvalidInput;  // a server reads the current table
    // no SRS until the new table is complete: a failure below leaves none
    srs_n_ = 0;
    srs_host_.clear();
    ZK_ALLOC(pts_in_, (uint64_t)n * 64);
    uint64_t stride = 0;
    int32_t rc = srs_alloc(n, &stride);
    if (rc != kOk) return rc;
    rc = stage_upload(pts_in_.p, points, (uint64_t)n * 64);
    if (rc != kOk) return rc;
    launch_build_table(CurveId::Bn254, pts_in_.p, srs_.p, n, stream_);
    rc = srs_finish(n, stride);
    if (rc != kOk) return rc;
    srs_n_ = n;
    srs_host_.assign(points, points + (uint64_t)n * 64);  // the caller's own bytes
    memcpy(srs_first_, points, 64);
    return kOk;
}

// SRS file (Sources/zkMetal/KZG/SRSFileFormat.swift:92-160): "ptau", u32
// version 1, u32 section count, then sections {u32 type, u64 length, data}.
// Section 1: u32 curve id (1 = BN254), u32 power, u32 ceremony power;
// section 2: G1 powers, 64 B per point (x, y big-endian standard form).
// Other sections (G2 powers, contributions) are skipped.
int32_t Engine::srs_load_ptau(const uint8_t* data, uint64_t len, uint32_t max_points,
                              uint32_t* n_loaded) {
    if (!data || len < 12) return kInvalidInput;
    if (servers_) return kInvalidInput;  // a server reads the current table
    auto rd32 = [&](uint64_t off) {
        return (uint32_t)data[off] | ((uint32_t)data[off + 1] << 8) | ((uint32_t)data[off + 2] << 16) |
               ((uint32_t)data[off + 3] << 24);
    };
    if (memcmp(data, "ptau", 4) != 0 || rd32(4) != 1) return kInvalidInput;
    const uint32_t nsec = rd32(8);
    uint64_t off = 12;
    const uint8_t* g1 = nullptr;
    uint64_t g1_len = 0;
    bool bn254 = false;
    for (uint32_t s = 0; s < nsec; s++) {
        if (off + 12 > len) return kInvalidInput;
        const uint32_t type = rd32(off);
        const uint64_t slen = (uint64_t)rd32(off + 4) | ((uint64_t)rd32(off + 8) << 32);
        off += 12;
        if (slen > len - off) return kInvalidInput;
        if (type == 1) {
            if (slen < 12) return kInvalidInput;
            bn254 = rd32(off) == 1;
        } else if (type == 2) {
            g1 = data + off;
            g1_len = slen;
        }
        off += slen;
    }
    if (!bn254 || !g1) return kInvalidInput;
    uint64_t n = g1_len / 64;
    if (max_points) n = std::min<uint64_t>(n, max_points);
    if (n == 0 || n >= (1ull << 31)) return kInvalidInput;
    srs_n_ = 0;
    ZK_ALLOC(pts_in_, n * 64);
    uint64_t stride = 0;
    {
        const int32_t rc = srs_alloc(n, &stride);
        if (rc != kOk) return rc;
    }
    ZK_ALLOC(err_, 16);
    {
        const int32_t rc = stage_upload(pts_in_.p, g1, n * 64);
        if (rc != kOk) return rc;
    }
    launch_fill_zero(err_.p, 16, stream_);
    launch_import_be(pts_in_.p, n, srs_.p, (uint32_t*)err_.p, stream_);
    uint32_t bad = 0;
    ZK_TRY(hipMemcpyAsync(&bad, err_.p, 4, hipMemcpyDeviceToHost, stream_));
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipGetLastError());
    if (bad) return kInvalidInput;  // not canonical or not on the curve
    {
        const int32_t rc = srs_finish(n, stride);
        if (rc != kOk) return rc;
    }
    srs_n_ = (uint32_t)n;
    const int32_t rc = srs_set_first();
    if (rc != kOk) return rc;
    if (n_loaded) *n_loaded = srs_n_;
    return kOk;
}

// b * 2^(8w) * G for w < 32, b < 256, affine Montgomery (entry 0 = (0,0))
// gtab[w*256 + b] = b * 2^(8w) * G, affine in the 2^256 Montgomery form of
// the field H works in, entry 0 = (0,0); uploaded into dst
template <class H>
static int32_t build_gtab(const zkh::Fp& gx, const zkh::Fp& gy, DevBuf& dst) {
    std::vector<zkh::Xyzz> pts(32 * 256);
    zkh::Xyzz base = H::from_affine(gx, gy);
    for (int w = 0; w < 32; w++) {
        zkh::Xyzz acc = H::identity();
        pts[w * 256] = acc;
        for (int b = 1; b < 256; b++) {
            acc = H::xadd(acc, base);
            pts[w * 256 + b] = acc;
        }
        base = H::xadd(acc, base);  // 256 * base
    }
    std::vector<zkh::Fp> xs, ys;
    H::batch_to_affine(pts, xs, ys);
    std::vector<uint8_t> blob(32 * 256 * 64);
    for (size_t i = 0; i < pts.size(); i++) {
        memcpy(&blob[i * 64], xs[i].v, 32);
        memcpy(&blob[i * 64 + 32], ys[i].v, 32);
    }
    // dst stays empty (not built) unless the upload completes
    if (dst.ensure(blob.size()) != hipSuccess) return kAllocFailed;
    if (hipMemcpy(dst.p, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
        dst.release();
        return kGpuError;
    }
    return kOk;
}

// BN254 G1 generator (1, 2)
int32_t Engine::ensure_gtab() {
    if (gtab_.p) return kOk;
    using H = zkh::HP;
    return build_gtab<H>(H::mul(zkh::Fp{{1, 0, 0, 0}}, H::r2()),
                         H::mul(zkh::Fp{{2, 0, 0, 0}}, H::r2()), gtab_);
}

// Grumpkin generator (1, sqrt(-16)) with the y of the Aztec/Noir
// specification (Sources/zkMetal/Curve/GrumpkinCurve.swift:215-222)
int32_t Engine::ensure_gtab_grumpkin() {
    if (gtab_gk_.p) return kOk;
    using H = zkh::HR;
    const zkh::Fp gy{{0x833fc48d823f272cULL, 0x2d270d45f1181294ULL, 0xcf135e7506a45d63ULL, 0x2ULL}};
    return build_gtab<H>(H::mul(zkh::Fp{{1, 0, 0, 0}}, H::r2()), H::mul(gy, H::r2()), gtab_gk_);
}

// BN254 G2 generator table: gtab[w*256 + b] = b * 2^(8w) * G2, affine
// 128-byte points (x.c0, x.c1, y.c0, y.c1 in the 2^256 form; entry 0 all
// zero), G2 the reference's generator (BN254Pairing.swift:389-396); one
// inversion for all entries (prefix products)
int32_t Engine::ensure_gtab_g2() {
    if (gtab_g2_.p) return kOk;
    using G = zkh::HostG2;
    using H = zkh::HP;
    auto mont = [](zkh::Fp v) { return H::mul(v, H::r2()); };
    const zkh::Fp2 gx{mont({{0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL}}),
                      mont({{0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL}})};
    const zkh::Fp2 gy{mont({{0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL}}),
                      mont({{0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL}})};
    std::vector<zkh::Xyzz2> pts(32 * 256);
    zkh::Xyzz2 base{gx, gy, G::one2(), G::one2()};
    for (int w = 0; w < 32; w++) {
        zkh::Xyzz2 acc = G::identity();
        pts[w * 256] = acc;
        for (int b = 1; b < 256; b++) {
            acc = G::xadd(acc, base);
            pts[w * 256 + b] = acc;
        }
        base = G::xadd(acc, base);  // 256 * base
    }
    // d_i = ZZ_i ZZZ_i (1 for the identity); 1/ZZ = ZZZ/d, 1/ZZZ = ZZ/d
    const size_t n = pts.size();
    std::vector<zkh::Fp2> pre(n + 1);
    pre[0] = G::one2();
    for (size_t i = 0; i < n; i++)
        pre[i + 1] = G::is_id(pts[i]) ? pre[i] : G::mul2(pre[i], G::mul2(pts[i].ZZ, pts[i].ZZZ));
    zkh::Fp2 acc = G::inv2(pre[n]);
    std::vector<uint8_t> blob(n * 128, 0);
    for (size_t i = n; i-- > 0;) {
        if (G::is_id(pts[i])) continue;
        const zkh::Fp2 dinv = G::mul2(acc, pre[i]);
        acc = G::mul2(acc, G::mul2(pts[i].ZZ, pts[i].ZZZ));
        const zkh::Fp2 x = G::mul2(pts[i].X, G::mul2(dinv, pts[i].ZZZ));
        const zkh::Fp2 y = G::mul2(pts[i].Y, G::mul2(dinv, pts[i].ZZ));
        const zkh::Fp* c[4] = {&x.c0, &x.c1, &y.c0, &y.c1};
        for (int k = 0; k < 4; k++) memcpy(&blob[i * 128 + 32 * k], c[k]->v, 32);
    }
    if (gtab_g2_.ensure(blob.size()) != hipSuccess) return kAllocFailed;
    if (hipMemcpy(gtab_g2_.p, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
        gtab_g2_.release();
        return kGpuError;
    }
    return kOk;
}

int32_t Engine::g2_points_generate(uint64_t seed, uint32_t n, void* d_out) {
    if (n == 0 || !d_out) return kInvalidInput;
    int rc = ensure_gtab_g2();
    if (rc != kOk) return rc;
    std::vector<uint8_t> ks((uint64_t)n * 32);
    zkh::gen_scalars(seed, n, 0, ks.data());
    ZK_ALLOC(tmp_, (uint64_t)n * 32);
    ZK_TRY(hipMemcpyAsync(tmp_.p, ks.data(), ks.size(), hipMemcpyHostToDevice, stream_));
    launch_gen_points_g2(tmp_.p, gtab_g2_.p, n, d_out, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipGetLastError());
    return kOk;
}

int32_t Engine::grumpkin_points_generate(uint64_t seed, uint32_t n, void* d_out) {
    if (n == 0 || !d_out) return kInvalidInput;
    int rc = ensure_gtab_grumpkin();
    if (rc != kOk) return rc;
    std::vector<uint8_t> ks((uint64_t)n * 32);
    zkh::gen_scalars(seed, n, 0, ks.data());
    ZK_ALLOC(tmp_, (uint64_t)n * 32);
    ZK_TRY(hipMemcpyAsync(tmp_.p, ks.data(), ks.size(), hipMemcpyHostToDevice, stream_));
    launch_gen_points_grumpkin(tmp_.p, gtab_gk_.p, n, d_out, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipGetLastError());
    return kOk;
}

int32_t Engine::srs_generate(uint64_t seed, uint32_t n) {
    if (n == 0) return kInvalidInput;
    if (servers_) return kInvalidInput;  // a server reads the current table
    srs_n_ = 0;  // no SRS until the new table is complete
    srs_host_.clear();
    int rc = ensure_gtab();
    if (rc != kOk) return rc;
    std::vector<uint8_t> ks((uint64_t)n * 32);
    zkh::gen_scalars(seed, n, 0, ks.data());
    ZK_ALLOC(tmp_, (uint64_t)n * 32);
    uint64_t stride = 0;
    rc = srs_alloc(n, &stride);
    if (rc != kOk) return rc;
    ZK_TRY(hipMemcpyAsync(tmp_.p, ks.data(), ks.size(), hipMemcpyHostToDevice, stream_));
    launch_gen_bases(tmp_.p, gtab_.p, n, srs_.p, stream_);
    rc = srs_finish(n, stride);
    if (rc != kOk) return rc;
    srs_n_ = n;
    return srs_set_first();
}

int32_t Engine::srs_download(uint8_t* points, uint32_t n) {
    if (!points || n > srs_n_) return kInvalidInput;
    if (n == 0) return kOk;
    // table entries 2i are the bases, in the kernels' internal form
    ZK_ALLOC(tmp_, (uint64_t)n * 64);
    launch_table_export(CurveId::Bn254, srs_.p, n, tmp_.p, stream_);
    ZK_TRY(hipMemcpyAsync(points, tmp_.p, (uint64_t)n * 64, hipMemcpyDeviceToHost, stream_));
    ZK_TRY(hipStreamSynchronize(stream_));
    return kOk;
}

int32_t Engine::msm_resident(const void* scalars, bool on_device, const uint32_t* counts,
                             uint32_t nmsm, uint8_t* out96) {
    if (srs_n_ == 0) return kInvalidInput;
    return msm_table(srs_.p, srs_n_, scalars, on_device, counts, nmsm, out96);
}

int32_t Engine::kzg_commit(const void* coeffs, bool on_device, const uint32_t* counts,
                           uint32_t npoly, uint8_t* out96) {
    if (srs_n_ == 0 || !coeffs || !counts) return kInvalidInput;
    if (on_device) return msm_table_parts(srs_.p, srs_n_, nullptr, coeffs, counts, npoly, out96, 1);
    std::vector<const uint8_t*> parts(npoly);
    uint64_t off = 0;
    for (uint32_t j = 0; j < npoly; j++) {
        parts[j] = (const uint8_t*)coeffs + off * 32;
        off += counts[j];
    }
    return msm_table_parts(srs_.p, srs_n_, parts.data(), nullptr, counts, npoly, out96, 1);
}

int32_t Engine::msm_table(const void* table, uint32_t table_n, const void* scalars,
                          bool on_device, const uint32_t* counts, uint32_t nmsm, uint8_t* out96) {
    if (!scalars || !counts) return kInvalidInput;
    if (on_device) return msm_table_parts(table, table_n, nullptr, scalars, counts, nmsm, out96);
    // host scalars, concatenated: one part per MSM
    std::vector<const uint8_t*> parts(nmsm);
    uint64_t off = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        parts[j] = (const uint8_t*)scalars + off * 32;
        off += counts[j];
    }
    return msm_table_parts(table, table_n, parts.data(), nullptr, counts, nmsm, out96);
}

int32_t Engine::msm_table_parts(const void* table, uint32_t table_n,
                                const uint8_t* const* host_parts, const void* d_scalars,
                                const uint32_t* counts, uint32_t nmsm, uint8_t* out96,
                                uint32_t form, uint64_t shift_stride) {
    if (!table || !counts || !out96 || nmsm == 0) return kInvalidInput;
    if (!host_parts && !d_scalars) return kInvalidInput;
    std::vector<uint32_t> ns, idx;
    std::vector<uint64_t> t0;
    uint64_t total = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        if (counts[j] > table_n) return kInvalidInput;
        if (counts[j] == 0) {
            memset(out96 + (uint64_t)j * 96, 0, 96);
            continue;
        }
        if (host_parts && !host_parts[j]) return kInvalidInput;
        ns.push_back(counts[j]);
        t0.push_back(0);  // every MSM uses a prefix of the table
        idx.push_back(j);
        total += counts[j];
    }
    if (ns.empty()) return kOk;
    if (total >= (1ull << 30)) return kInvalidInput;
    Plan& pl = plan_;
    int rc = make_plan(ns.data(), t0.data(), (uint32_t)ns.size(), window_bits_for(ns).data(), pl, form, acc_cap_,
                       CurveId::Bn254, parts_req_, plan_shift(table, shift_stride), kShiftC,
                       shift_floor((uint32_t)ns.size()), schedule((uint32_t)ns.size()));
    if (rc != kOk) return rc;
    const void* d_sc = d_scalars;
    if (host_parts) {
        ZK_ALLOC(scalars_, total * 32);
        uint64_t off = 0;
        for (uint32_t k = 0; k < idx.size(); k++) {
            const uint64_t bytes = (uint64_t)ns[k] * 32;
            rc = stage_upload((uint8_t*)scalars_.p + off, host_parts[idx[k]], bytes);
            if (rc != kOk) return rc;
            off += bytes;
        }
        d_sc = scalars_.p;
    }
    std::vector<uint8_t> res(ns.size() * 96);
    rc = run(pl, d_sc, table, table_n, res.data());
    if (rc != kOk) return rc;
    for (size_t k = 0; k < idx.size(); k++) memcpy(out96 + (uint64_t)idx[k] * 96, res.data() + k * 96, 96);
    return kOk;
}

uint32_t Engine::window_bit(uint32_t n, bool batch, CurveId cv) const {
    uint32_t c = c_override_;
    if (!c && cv == CurveId::Bn254) {
        uint32_t cls = 0;
        while ((1ull << cls) < n) cls++;  // ceil(log2 n)
        cls = std::min<uint32_t>(cls, 32);
        if (batch) c = ctab_batch_[cls];
        if (!c) c = ctab_[cls];
    }
    return c;  // 0 = cost-model default
}

std::vector<uint32_t> Engine::window_bits_for(const std::vector<uint32_t>& ns, CurveId cv) const {
    std::vector<uint32_t> cs(ns.size());
    for (size_t j = 0; j < ns.size(); j++) cs[j] = window_bit(ns[j], ns.size() > 1, cv);
    return cs;
}

int32_t Engine::set_window_table(const uint32_t* c_by_log2n, uint32_t len, bool batch) {
    uint32_t t[33];
    for (uint32_t i = 0; i <= 32; i++) {
        const uint32_t c = (c_by_log2n && i < len) ? c_by_log2n[i] : 0;
        if (c != 0 && (c < 2 || c > 16)) return kInvalidInput;
        t[i] = c;
    }
    memcpy(batch ? ctab_batch_ : ctab_, t, sizeof(t));
    return kOk;
}

int32_t Engine::set_schedule_table(const uint32_t* parts, const uint32_t* per, uint32_t len, bool batch) {
    Schedule t{};
    for (uint32_t i = 0; i <= 32; i++) {
        const uint32_t p = (parts && i < len) ? parts[i] : 0;
        const uint32_t g = (per && i < len) ? per[i] : 0;
        // per: a power of two from 4 (a column of up to 256 buckets then has
        // at most 64 lanes, one wave's shuffle tree) to 64
        if (p > kMaxParts || (g != 0 && (g < 4 || g > 64 || (g & (g - 1)) != 0))) return kInvalidInput;
        t.parts[i] = (uint8_t)p;
        t.per[i] = (uint8_t)g;
    }
    sched_[batch ? 1 : 0] = t;
    return kOk;
}

// ------------------------------------------------------ no-allocation path
// The device buffers of zkmsm_msm_batch_resident_device_noalloc: a fixed
// array in the library's device data, one copy per device (the runtime puts
// it on each device that loads the library; it stays until the process
// ends).  In host code `scratch` names a host-side stand-in, not device
// memory: host code uses only the address hipGetSymbolAddress returned for
// the current device (kept per device in g_scratch, copied into each engine).
static __device__ __align__(256) uint8_t scratch[Engine::kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};
// one device's array serves one call at a time
std::mutex g_noalloc_mu[kMaxDevices];
constexpr size_t kNoAllocHostBytes = 512u << 10;
}  // namespace

// at engine creation, with the engine's device current: the device address
// of this device's array (looked up once per device), and host capacity for
// the plan and the staging, so that the calls themselves allocate nothing
int32_t Engine::noalloc_setup() {
    if (dev_ < 0 || dev_ >= kMaxDevices) return kGpuError;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[dev_]) {
            void* ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
                (void)hipGetLastError();
                return kGpuError;
            }
            g_scratch[dev_] = (uint8_t*)ptr;
        }
        scratch_ = g_scratch[dev_];
    }
    na_plan_.msms.reserve(kNoAllocMaxMsms);
    na_plan_.msm_scalar0.reserve(kNoAllocMaxMsms + 1);
    na_plan_.msm_dtile0.reserve(kNoAllocMaxMsms + 1);
    na_plan_.segs.reserve(kNoAllocMaxSegs);
    na_plan_.seg_lane0.reserve(kNoAllocMaxSegs + 1);
    na_plan_.seg_grp0.reserve(kNoAllocMaxSegs + 1);
    na_plan_.seg_rcw0.reserve(kNoAllocMaxSegs + 1);
    na_plan_.tile_seg.reserve(4 * kNoAllocMaxSegs);
    na_plan_.chunk_seg.reserve(2 * kNoAllocMaxSegs);
    na_host_.reserve(kNoAllocHostBytes);
    return kOk;
}

int32_t Engine::msm_resident_noalloc(const void* d_scalars, const uint32_t* counts, uint32_t nmsm,
                                     uint8_t* out96) {
    if (!scratch_) return kGpuError;  // the array's address was not found at creation
    if (!d_scalars || !counts || !out96 || nmsm == 0 || srs_n_ == 0) return kInvalidInput;
    if (nmsm > kNoAllocMaxMsms) return kInvalidInput;
    uint32_t ns[kNoAllocMaxMsms], cs[kNoAllocMaxMsms], idx[kNoAllocMaxMsms];
    uint64_t t0[kNoAllocMaxMsms];
    uint32_t k = 0;
    uint64_t total = 0;
    for (uint32_t j = 0; j < nmsm; j++) {
        if (counts[j] > srs_n_) return kInvalidInput;
        total += counts[j];
        if (counts[j] == 0) continue;
        ns[k] = counts[j];
        t0[k] = 0;
        idx[k++] = j;
    }
    if (total > kNoAllocMaxPoints) return kInvalidInput;
    for (uint32_t j = 0; j < nmsm; j++)
        if (counts[j] == 0) memset(out96 + (uint64_t)j * 96, 0, 96);
    if (k == 0) return kOk;
    // the plan's array sizes, before planning: within the reserved capacity
    uint64_t nseg = 0, tiles = 0, chunks = 0;
    for (uint32_t i = 0; i < k; i++) {
        cs[i] = window_bit(ns[i], k > 1, CurveId::Bn254);
        const uint32_t c = cs[i] ? cs[i] : default_window_bits(ns[i]);
        if (c < 2 || c > 16) return kInvalidInput;
        const uint64_t w = (128 + c - 1) / c, nt = (2ull * ns[i] + kSortTile - 1) / kSortTile;
        nseg += w;
        tiles += w * nt;
        chunks += w * ((256 * nt + kScanChunk - 1) / kScanChunk);
    }
    if (nseg > na_plan_.segs.capacity() || tiles > na_plan_.tile_seg.capacity() ||
        chunks > na_plan_.chunk_seg.capacity())
        return kInvalidInput;
    // at most 2^15 accumulation lanes: a quarter of the device's resident
    // lanes, ample for <= 2^16 points, and it bounds the merge records (two
    // per lane), the largest buffer of a small call
    int rc = make_plan(ns, t0, k, cs, na_plan_, 0, std::min<uint64_t>(acc_cap_, 1u << 15), CurveId::Bn254, 0,
                       plan_shift(srs_.p, kOwnTable), kShiftC, shift_mode_ == kShiftAll ? 1u : 0u);
    if (rc != kOk) return rc;
    const Plan& pl = na_plan_;
    const WsSizes z = ws_sizes(pl, true);
    const uint64_t desc_host = (z.desc + 255) & ~uint64_t(255);
    const uint64_t res_host = (uint64_t)k * 96;
    if (z.device > kNoAllocScratchBytes || desc_host + z.host + res_host > na_host_.capacity())
        return kInvalidInput;  // the call needs more than the array holds
    // carve the call's buffers out of the array, 256-byte aligned, in the
    // order ws_sizes counts them
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        void* p = scratch_ + off;
        off += (bytes + 255) & ~uint64_t(255);
        return p;
    };
    Workspace ws;
    ws.desc = (uint8_t*)take(z.desc);
    ws.digits = take(z.digits);
    ws.keysA = take(z.keys);
    ws.valsA = take(z.vals);
    ws.keysB = take(z.keys);
    ws.valsB = take(z.vals);
    ws.hist = take(z.hist);
    ws.part = take(z.part);
    ws.segvalid = take(z.segvalid);
    ws.buckets = take(z.buckets);
    ws.recAk = take(z.rec_keys);
    ws.recAp = take(z.rec_pts);
    ws.recBk = take(z.rec_keys);
    ws.rc = take(z.rc);
    ws.bits = take(z.bits);
    ws.err = take(z.err);
    ws.wins = take(z.wins);
    ws.res = take(z.res);
    na_host_.resize(desc_host + z.host + res_host);  // within the reserved capacity
    ws.hdesc = na_host_.data();
    ws.hw = na_host_.data() + desc_host;
    uint8_t* res = ws.hw + z.host;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu[dev_]);
        rc = run_ws(pl, ws, z, true, kGraphOff, d_scalars, srs_.p, srs_n_, res);
    }
    if (rc != kOk) return rc;
    for (uint32_t i = 0; i < k; i++) memcpy(out96 + (uint64_t)idx[i] * 96, res + (uint64_t)i * 96, 96);
    return kOk;
}

// ------------------------------------------------------------ unit probes

int32_t Engine::probe_copy(uint64_t bytes, double* ms_out) {
    if (bytes < 16 || !ms_out) return kInvalidInput;
    DevBuf a, b;
    ZK_ALLOC(a, bytes);
    ZK_ALLOC(b, bytes);
    launch_fill_zero(a.p, bytes, stream_);
    launch_copy(a.p, b.p, bytes, stream_);  // warm-up
    ZK_TRY(hipEventRecord(ev_[0], stream_));
    launch_copy(a.p, b.p, bytes, stream_);
    ZK_TRY(hipEventRecord(ev_[1], stream_));
    ZK_TRY(hipEventSynchronize(ev_[1]));
    float ms = 0;
    ZK_TRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    *ms_out = ms;
    return kOk;
}

int32_t Engine::probe_rates(double* out6) {
    if (!out6) return kInvalidInput;
    hipDeviceProp_t prop;
    ZK_TRY(hipGetDeviceProperties(&prop, dev_));
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 8;
    const uint64_t threads = (uint64_t)blocks * 256;
    DevBuf o;
    ZK_ALLOC(o, threads * 8);
    // modes 0-2, 4: throughput at full occupancy; 3, 5: one 256-thread block
    // of dependent multiplies (one wave per SIMD: a single lane's chain)
    const uint32_t iters[6] = {4096, 4096, 64, 256, 64, 256};
    const double per_iter[6] = {8.0, 8.0, 4.0, 1.0, 4.0, 1.0};
    for (int mode = 0; mode < 6; mode++) {
        const bool chain = mode == 3 || mode == 5;
        const uint32_t nb = chain ? 1 : blocks;
        launch_probe(mode, nb, 16, (uint64_t*)o.p, stream_);  // warm-up
        ZK_TRY(hipEventRecord(ev_[0], stream_));
        launch_probe(mode, nb, iters[mode], (uint64_t*)o.p, stream_);
        ZK_TRY(hipEventRecord(ev_[1], stream_));
        ZK_TRY(hipEventSynchronize(ev_[1]));
        float ms = 0;
        ZK_TRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
        if (!chain) out6[mode] = (double)threads * iters[mode] * per_iter[mode] / (ms * 1e-3);
        else out6[mode] = ms * 1e6 / iters[mode];  // ns per dependent multiply
    }
    return kOk;
}

int32_t Engine::test_fp(int op, const uint8_t* a, const uint8_t* b, uint8_t* out, uint32_t n) {
    if (!a || !b || !out || n == 0) return kInvalidInput;
    DevBuf da, dbb, dout;
    ZK_ALLOC(da, (uint64_t)n * 32);
    ZK_ALLOC(dbb, (uint64_t)n * 32);
    ZK_ALLOC(dout, (uint64_t)n * 32);
    ZK_TRY(hipMemcpy(da.p, a, (uint64_t)n * 32, hipMemcpyHostToDevice));
    ZK_TRY(hipMemcpy(dbb.p, b, (uint64_t)n * 32, hipMemcpyHostToDevice));
    launch_test_fp(op, da.p, dbb.p, dout.p, n, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipMemcpy(out, dout.p, (uint64_t)n * 32, hipMemcpyDeviceToHost));
    return kOk;
}

int32_t Engine::test_xyzz(int op, const uint8_t* p, const uint8_t* q, uint8_t* out, uint32_t n) {
    if (!p || !q || !out || n == 0) return kInvalidInput;
    DevBuf dp, dq, dout;
    ZK_ALLOC(dp, (uint64_t)n * 128);
    ZK_ALLOC(dq, (uint64_t)n * 128);
    ZK_ALLOC(dout, (uint64_t)n * 128);
    ZK_TRY(hipMemcpy(dp.p, p, (uint64_t)n * 128, hipMemcpyHostToDevice));
    ZK_TRY(hipMemcpy(dq.p, q, (uint64_t)n * 128, hipMemcpyHostToDevice));
    launch_test_xyzz(op, dp.p, dq.p, dout.p, n, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipMemcpy(out, dout.p, (uint64_t)n * 128, hipMemcpyDeviceToHost));
    return kOk;
}

int32_t Engine::test_g2(int op, const uint8_t* p, const uint8_t* q, uint8_t* out, uint32_t n) {
    if (!p || !q || !out || n == 0 || op < 0 || op > 6) return kInvalidInput;
    DevBuf dp, dq, dout;
    const uint64_t bytes = (uint64_t)n * 256;
    ZK_ALLOC(dp, bytes);
    ZK_ALLOC(dq, bytes);
    ZK_ALLOC(dout, bytes);
    ZK_TRY(hipMemcpy(dp.p, p, bytes, hipMemcpyHostToDevice));
    ZK_TRY(hipMemcpy(dq.p, q, bytes, hipMemcpyHostToDevice));
    launch_test_g2(op, dp.p, dq.p, dout.p, n, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipMemcpy(out, dout.p, bytes, hipMemcpyDeviceToHost));
    return kOk;
}

int32_t Engine::test_glv(const uint8_t* scalars, uint8_t* out, uint32_t n) {
    if (!scalars || !out || n == 0) return kInvalidInput;
    DevBuf ds, dout;
    ZK_ALLOC(ds, (uint64_t)n * 32);
    ZK_ALLOC(dout, (uint64_t)n * 48);
    ZK_TRY(hipMemcpy(ds.p, scalars, (uint64_t)n * 32, hipMemcpyHostToDevice));
    launch_test_glv(ds.p, dout.p, n, stream_);
    ZK_TRY(hipStreamSynchronize(stream_));
    ZK_TRY(hipMemcpy(out, dout.p, (uint64_t)n * 48, hipMemcpyDeviceToHost));
    return kOk;
}

}  // namespace zk
