// MSM engine: device workspace, batch planner, launch sequence and the host
// tail.  One engine owns one HIP stream on one device; calls are serialised
// by the engine mutex (the reference engine is not re-entrant at all,
// Sources/zkMetal/MSM/MSMEngine.swift:232-277).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "host_curve.hpp"
#include "plan.hpp"

namespace zk {

enum Status : int32_t {
    kOk = 0,
    kNoGpu = -1,
    kInvalidInput = -2,
    kGpuError = -3,
    kAllocFailed = -4,
};

// device bytes held by one engine's buffers, and an optional cap on them
struct WsAccount {
    uint64_t limit = 0;  // 0: no cap
    uint64_t used = 0;
};

// grow-only device buffer; with an account, a growth that would take the
// account past its limit fails (hipErrorOutOfMemory) and keeps the old buffer
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    WsAccount* acct = nullptr;
    hipError_t ensure(size_t bytes);
    void release();
    ~DevBuf() { release(); }
};

// grow-only pinned host buffer
struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    ~HostBuf() { release(); }
};

enum Stage {
    kStDecompose = 0,
    kStSort,
    kStAccumulate,
    kStMerge,
    kStReduce,
    kStCopyOut,
    kNumStages
};

struct Timing {
    float stage_ms[kNumStages];  // device time per stage (HIP events on the engine stream)
    float device_ms;             // first to last event
    float host_tail_ms;          // window combine + normalisation on the host
    float device_tail_ms;        // the same on the device (0 when the host ran it)
    float wall_ms;               // whole call, host clock
    uint32_t nmsm, window_bits_first, T, G, parts;
    uint64_t entries, buckets, lanes;
    uint64_t valid_entries;  // sorted entries with a nonzero digit (mixed additions done)
};

// A call's segments split into consecutive parts: the sort of part p + 1
// (its own stream) runs while part p accumulates.  Ranges are [lo, hi).
constexpr uint32_t kMaxParts = 8;
struct PlanPart {
    uint32_t seg_lo, seg_hi, tile_lo, tile_hi, chunk_lo, chunk_hi;
    uint32_t T;  // entries per accumulation lane in this part
    uint64_t lane_lo, lane_hi, rcw_lo, rcw_hi, entries;
};

struct Plan {
    std::vector<MsmDesc> msms;
    std::vector<uint64_t> msm_scalar0;  // nmsm + 1
    // nmsm + 1: first decomposition block of each MSM; a block decomposes
    // the kSortTile / 2 scalars of one first-pass sort tile of every window
    std::vector<uint64_t> msm_dtile0;
    std::vector<SegDesc> segs;
    std::vector<uint64_t> seg_lane0;  // nseg + 1
    std::vector<uint64_t> seg_grp0;   // nseg + 1: first row/column sum of the segment
    std::vector<uint64_t> seg_rcw0;   // nseg + 1: first row/column wave of the segment
    std::vector<uint32_t> tile_seg;
    std::vector<uint32_t> chunk_seg;
    uint64_t scalars = 0, entries = 0, buckets = 0, lanes = 0, groups = 0, rc_waves = 0;
    uint32_t tiles = 0, chunks = 0, T = 0, dtiles = 0;  // T: the first part's
    uint32_t nparts = 1;
    PlanPart parts[kMaxParts];
    CurveId curve = CurveId::Bn254;
    uint32_t G = 8;  // buckets per bucket-reduction lane (a power of 2, 8..32)
};

// Device buffers of one call (and the host staging it copies through).
struct Workspace {
    uint8_t* desc = nullptr;  // packed descriptor tables
    void *digits = nullptr, *keysA = nullptr, *valsA = nullptr, *keysB = nullptr, *valsB = nullptr;
    void *hist = nullptr, *part = nullptr, *segvalid = nullptr, *buckets = nullptr;
    void *recAk = nullptr, *recAp = nullptr, *recBk = nullptr, *rc = nullptr, *bits = nullptr;
    void *err = nullptr, *wins = nullptr, *res = nullptr;
    uint8_t* hdesc = nullptr;  // host: descriptors before upload
    uint8_t* hw = nullptr;     // host: copy-back area
};

// bytes of each buffer of a planned call; device = all device buffers, each
// rounded up to 256 bytes; host = the copy-back area
struct WsSizes {
    uint64_t desc, digits, keys, vals, hist, part, segvalid, buckets, rec_keys, rec_pts, rc, bits, err,
        wins, res, back, host, device;
};
WsSizes ws_sizes(const Plan& pl, bool dev_tail);

// window width for an MSM of n points (cost model; overridable)
uint32_t default_window_bits(uint64_t n, CurveId cv = CurveId::Bn254);

// Launch schedule by size class ceil(log2 entries) of a call (its sorted
// entries before the zero digits are dropped): parts[k] the pipeline parts
// (1..kMaxParts), per[k] the buckets each bucket-reduction lane sums before
// the shuffle tree (4..64, a power of two); 0 keeps the built-in rule.
struct Schedule {
    uint8_t parts[33];
    uint8_t per[33];
};

// build a plan; ns[j] > 0 for all j; table0[j] = first table entry of MSM j;
// nparts: how many parts to split the segments into (0: the schedule's entry
// for the call's size class, else by size, see default_parts; at most
// kMaxParts and the segment count); sched (nullable): the size-class
// schedule, which also sets the reduction's buckets per lane; shift_stride > 0:
// the table also holds window-shifted copies for window width shift_c, copy w
// at entry w * shift_stride (a multiple of kSortTile), and MSMs run
// window-shifted at width shift_c when shifted_pays or (shift_floor > 0) they
// have at least shift_floor points
int make_plan(const uint32_t* ns, const uint64_t* table0, uint32_t nmsm,
              const uint32_t* cs /* nullable */, Plan& pl, uint32_t form = 0,
              uint64_t lane_cap = 1u << 17, CurveId cv = CurveId::Bn254, uint32_t nparts = 0,
              uint64_t shift_stride = 0, uint32_t shift_c = 0, uint32_t shift_floor = 0,
              const Schedule* sched = nullptr);
uint32_t default_parts(uint64_t entries);
bool shifted_pays(uint64_t n, uint32_t c, uint32_t c_s);

class Engine {
  public:
    static int32_t create(Engine** out);
    ~Engine();
    // Shared ownership: a request server attached to this engine's SRS holds
    // a reference, so destroying the caller's handle first leaves the table
    // alive until the last server is gone.  release() deletes at zero.
    void retain() { refs_.fetch_add(1, std::memory_order_relaxed); }
    static void release(Engine* e) {
        if (e && e->refs_.fetch_sub(1, std::memory_order_acq_rel) == 1) delete e;
    }
    // servers reading the resident table (engine mutex held by the caller):
    // while any is attached, the srs_* calls that replace the table refuse
    void attach_server() { servers_++; }
    void detach_server() { servers_--; }
    uint32_t attached_servers() const { return servers_; }

    // host buffers in, host results out (96 B Jacobian X|Y|Z per MSM)
    int32_t msm_host(const uint8_t* points, const uint8_t* scalars, uint32_t n, uint8_t* out96);
    // cv = CurveId::Grumpkin: points are Grumpkin affine points (x, y in Fr,
    // Montgomery form), scalars are reduced mod p, results are Grumpkin points
    int32_t msm_batch_host(const uint8_t* points, const uint8_t* scalars,
                           const uint32_t* counts, uint32_t nmsm, uint8_t* out96,
                           CurveId cv = CurveId::Bn254);

    // k MSMs over one base set of n points (scalar set j at scalars + j*n*32):
    // the table is built once and the k MSMs run as one batched launch
    int32_t multi_msm_host(const uint8_t* points, uint32_t n, const uint8_t* scalars, uint32_t k,
                           uint8_t* out96);

    // bases and scalars already in device memory (plain MSMs, table built per call)
    int32_t msm_device(const void* d_points, const void* d_scalars, const uint32_t* counts,
                       uint32_t nmsm, uint8_t* out96, CurveId cv = CurveId::Bn254);
    // n Grumpkin points k_i * G (k_i = gen_scalars(seed, n, 0)) into a device
    // buffer of n * 64 bytes
    int32_t grumpkin_points_generate(uint64_t seed, uint32_t n, void* d_out);
    // n BN254 G2 points k_i * G2 (128 B each) into a device buffer
    int32_t g2_points_generate(uint64_t seed, uint32_t n, void* d_out);

    // device-resident base table (SRS)
    int32_t srs_upload(const uint8_t* points, uint32_t n);
    int32_t srs_generate(uint64_t seed, uint32_t n);
    int32_t srs_download(uint8_t* points, uint32_t n);
    // G1 powers of an SRS file image (see engine.hip); max_points 0 = all
    int32_t srs_load_ptau(const uint8_t* data, uint64_t len, uint32_t max_points, uint32_t* n_loaded);
    uint32_t srs_size() const { return srs_n_; }
    // host-point calls whose bases are byte-for-byte prefixes of the resident
    // SRS run on the resident table (no base upload, no table build); on by
    // default
    void set_srs_detect(bool on) { srs_detect_ = on; }
    uint64_t srs_detect_hits() const { return srs_hits_; }

    // window-shifted resident tables: with mode kShiftAuto (default) or
    // kShiftAll an SRS set afterwards also gets the copies 2^(kShiftC w) P_i
    // of every window w (memory: the table times ceil(128 / kShiftC)); MSMs
    // on it share one bucket set across windows when shifted_pays (auto) or
    // always (all).  kShiftOff: plain plans, and the next SRS is plain.
    // Falls back to the plain table when the copies do not fit.  Results are
    // the same bytes in every mode.
    static constexpr uint32_t kShiftC = 16;
    enum ShiftMode { kShiftOff = 0, kShiftAuto = 1, kShiftAll = 2 };
    int32_t set_window_shift(int32_t mode) {
        if (mode < kShiftOff || mode > kShiftAll) return kInvalidInput;
        shift_mode_ = mode;
        return kOk;
    }
    // kShiftC when the resident table holds the shifted copies, else 0
    uint32_t srs_window_shift() const { return srs_shift_ ? kShiftC : 0; }
    uint64_t srs_shift_stride() const { return srs_shift_; }

    // MSMs over prefixes of the resident table; scalars on host or device
    int32_t msm_resident(const void* scalars, bool scalars_on_device, const uint32_t* counts,
                         uint32_t nmsm, uint8_t* out96);
    // same over any device table [P_0, phi(P_0), ...] of table_n bases (shared SRS)
    int32_t msm_table(const void* table, uint32_t table_n, const void* scalars,
                      bool scalars_on_device, const uint32_t* counts, uint32_t nmsm,
                      uint8_t* out96);
    // scalars given per MSM: host_parts[j] (n_j x 32 B, host) or, when
    // host_parts is null, one device buffer holding all MSMs' scalars
    // shift_stride: the table's window-shifted copies (srs_shift_stride() of
    // the engine that owns it; 0 = none; kOwnTable = this engine's resident
    // table's when table is it)
    static constexpr uint64_t kOwnTable = ~0ull;
    int32_t msm_table_parts(const void* table, uint32_t table_n,
                            const uint8_t* const* host_parts, const void* d_scalars,
                            const uint32_t* counts, uint32_t nmsm, uint8_t* out96,
                            uint32_t form = 0, uint64_t shift_stride = kOwnTable);
    // KZG commitments over the resident table: coefficients are Fr elements in
    // Montgomery form (converted to standard form on the device)
    int32_t kzg_commit(const void* coeffs, bool on_device, const uint32_t* counts, uint32_t npoly,
                       uint8_t* out96);
    const void* srs_table() const { return srs_.p; }

    // per-size-class window widths: c_by_log2n[ceil(log2 n)], 0 = cost model
    // per-size-class window widths for single-MSM calls (batch = false) and
    // for MSMs inside batched calls (batch = true); 0 = cost-model default
    int32_t set_window_table(const uint32_t* c_by_log2n, uint32_t len, bool batch = false);
    // per-size-class launch schedule (class ceil(log2 entries) of the call):
    // pipeline parts (0..kMaxParts) and the bucket reduction's buckets per
    // lane (0 or a power of two in 4..64); 0 keeps the built-in rule.  batch:
    // the table for calls of several MSMs, else for calls of one
    int32_t set_schedule_table(const uint32_t* parts, const uint32_t* per, uint32_t len, bool batch);
    const Schedule* schedule(uint32_t nmsm) const { return &sched_[nmsm > 1 ? 1 : 0]; }
    void copy_tuning_from(const Engine& o) {
        memcpy(ctab_, o.ctab_, sizeof(ctab_));
        memcpy(ctab_batch_, o.ctab_batch_, sizeof(ctab_batch_));
        memcpy(sched_, o.sched_, sizeof(sched_));
        c_override_ = o.c_override_;
        shift_mode_ = o.shift_mode_;
    }
    // per-MSM window widths: the override, else the tuned tables (BN254 G1
    // only), else 0 = the cost model
    std::vector<uint32_t> window_bits_for(const std::vector<uint32_t>& ns,
                                          CurveId cv = CurveId::Bn254) const;
    uint32_t window_bit(uint32_t n, bool batch, CurveId cv) const;

    void set_window_bits(uint32_t c) { c_override_ = c; }
    // cap on the device memory this engine's buffers may grow to (resident
    // SRS included; 0 = none): a call that needs more returns kAllocFailed
    // and leaves the engine usable
    void set_workspace_limit(uint64_t bytes) { ws_.limit = bytes; }
    uint64_t workspace_bytes() const { return ws_.used; }
    // graph mode: the device launch sequence of a call is captured into a
    // hipGraph and replayed while calls keep the same shape and buffers
    // kGraphAuto (default): a call whose shape and buffers repeat the
    // previous call's is captured once and replayed from then on
    enum GraphMode : int32_t { kGraphOff = 0, kGraphOn = 1, kGraphAuto = 2 };
    int32_t set_graph_mode(int32_t mode) {
        if (mode < kGraphOff || mode > kGraphAuto) return kInvalidInput;
        graph_mode_ = mode;
        return kOk;
    }
    // parts a call's segments are split into for the sort / accumulation
    // overlap (0: by size); a measurement knob, results are the same
    void set_parts(uint32_t p) { parts_req_ = std::min<uint32_t>(p, kMaxParts); }
    // where a call's window combine and normalisation run: kTailAuto (the
    // device in graph mode and for calls of at least kDeviceTailMinMsms MSMs,
    // else the host), kTailHost, kTailDevice.  BN254 G2 always uses the host.
    enum TailMode : int32_t { kTailAuto = 0, kTailHost = 1, kTailDevice = 2 };
    static constexpr uint32_t kDeviceTailMinMsms = 16;
    bool device_tail(uint32_t nmsm, CurveId cv) const {
        return cv != CurveId::Bn254G2 &&
               (tail_mode_ == kTailDevice ||
                (tail_mode_ == kTailAuto && (graph_mode_ == kGraphOn || nmsm >= kDeviceTailMinMsms)));
    }
    int32_t set_tail_mode(int32_t m) {
        if (m < kTailAuto || m > kTailDevice) return kInvalidInput;
        tail_mode_ = m;
        return kOk;
    }
    uint64_t graph_captures() const { return graph_captures_; }
    const Timing& last_timing() const { return timing_; }
    int device() const { return dev_; }
    hipStream_t stream() const { return stream_; }
    std::mutex& mutex() { return mu_; }

    // primitive tests (host in / host out)
    int32_t test_fp(int op, const uint8_t* a, const uint8_t* b, uint8_t* out, uint32_t n);
    int32_t test_xyzz(int op, const uint8_t* p, const uint8_t* q, uint8_t* out, uint32_t n);
    int32_t test_g2(int op, const uint8_t* p, const uint8_t* q, uint8_t* out, uint32_t n);
    int32_t test_glv(const uint8_t* scalars, uint8_t* out, uint32_t n);
    // measured: [0] v_mad_u64_u32 products/s, [1] FIPS accumulate steps/s,
    // [2] field multiplies/s (all at full occupancy); [3] ns per dependent
    // field multiply in a single lane (one wave per SIMD)
    int32_t probe_rates(double* out6);
    // one launch of a plain copy kernel over `bytes` (HBM counter calibration)
    int32_t probe_copy(uint64_t bytes, double* ms_out);

    // run a planned batch: scalars and table already on the device
    int32_t run(const Plan& pl, const void* d_scalars, const void* d_table, uint64_t table_n,
                uint8_t* out96);

    // zkmsm_msm_batch_resident_device without any allocation: the call's
    // buffers come from a fixed device array of kNoAllocScratchBytes (one
    // per device, shared by every engine on it; calls on one device are
    // serialised), the host side from capacity reserved when the engine was
    // created.  At most kNoAllocMaxPoints points in all and kNoAllocMaxMsms
    // MSMs, and the call's buffers must fit the array: otherwise
    // kInvalidInput.  The window combine runs on the device.
    static constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
    static constexpr uint32_t kNoAllocMaxPoints = 1u << 16;
    static constexpr uint32_t kNoAllocMaxMsms = 64;
    static constexpr uint32_t kNoAllocMaxSegs = 1024;
    int32_t msm_resident_noalloc(const void* d_scalars, const uint32_t* counts, uint32_t nmsm,
                                 uint8_t* out96);

  private:
    Engine() = default;
    int32_t run_ws(const Plan& pl, const Workspace& ws, const WsSizes& z, bool dev_tail, int32_t graph_mode,
                   const void* d_scalars, const void* d_table, uint64_t table_n, uint8_t* out96);
    int32_t noalloc_setup();  // the device array's address, host capacity
    int32_t ensure_gtab();
    int32_t ensure_gtab_grumpkin();
    int32_t ensure_gtab_g2();

    int dev_ = 0;
    hipStream_t stream_ = nullptr;
    std::mutex mu_;
    std::atomic<int32_t> refs_{1};
    uint32_t servers_ = 0;
    WsAccount ws_;
    void account_buffers();  // ties every member DevBuf to ws_
    uint32_t c_override_ = 0;
    uint32_t ctab_[33] = {};
    uint32_t ctab_batch_[33] = {};
    Schedule sched_[2] = {};  // [0]: calls of one MSM, [1]: of several
    int32_t graph_mode_ = kGraphAuto;
    std::vector<uint64_t> last_key_;  // the previous call's shape key (kGraphAuto)
    int32_t tail_mode_ = kTailAuto;
    uint64_t acc_cap_ = 1u << 17;  // resident k_accumulate lanes on this device
    uint64_t acc_cap_g2_ = 1u << 16;  // the same for the G2 accumulation
    uint64_t lane_cap(CurveId cv) const { return cv == CurveId::Bn254G2 ? acc_cap_g2_ : acc_cap_; }
    hipGraphExec_t graph_exec_ = nullptr;
    std::vector<uint64_t> graph_key_;
    uint64_t graph_captures_ = 0;
    Timing timing_{};

    DevBuf desc_, digits_, keysA_, valsA_, keysB_, valsB_, hist_, part_, segvalid_, buckets_;
    DevBuf recAk_, recAp_, recBk_, rc_, winsum_, err_;  // recBk_: long-run list
    DevBuf wins_, res_;  // device tail: window sums, results
    DevBuf pts_in_, ptable_, scalars_, srs_, gtab_, gtab_gk_, gtab_g2_, tmp_;
    HostBuf hdesc_, hwin_;
    uint32_t srs_n_ = 0;
    int32_t shift_mode_ = kShiftAuto;
    Plan plan_;  // the device-input and resident calls' plan, its arrays reused from call to call
    uint64_t srs_shift_ = 0;  // entries between the resident table's window copies (0: none)
    // allocate the resident table for n bases; *stride = the window copies'
    // stride when they fit (built by srs_finish), else 0
    int32_t srs_alloc(uint64_t n, uint64_t* stride);
    // after the table's bases are written: the shifted copies, then wait
    int32_t srs_finish(uint64_t n, uint64_t stride);
    uint64_t plan_shift(const void* table, uint64_t shift_stride) const;
    // make_plan's shift_floor for a call of nmsm MSMs: every MSM in kShiftAll
    // mode; in a call whose window combine runs on the device, every MSM of
    // at least kShiftDeviceTailFloor points (a plain MSM's serial doubling
    // chain would set the combine's latency)
    static constexpr uint32_t kShiftDeviceTailFloor = 4096;
    uint32_t shift_floor(uint32_t nmsm) const;
    // host image of the resident bases (64 B each) for the prefix check:
    // srs_first_ is filled whenever an SRS is set, srs_host_ on the first
    // call whose first base matches it
    std::vector<uint8_t> srs_host_;
    uint8_t srs_first_[64] = {};
    bool srs_detect_ = true;
    uint64_t srs_hits_ = 0;
    int32_t srs_set_first();
    // cheap test: every point slice fits the SRS and starts with its first
    // base (and the host image for the full check is present)
    bool srs_candidate(const uint8_t* points, const uint32_t* counts, uint32_t nmsm);
    // exact test of a candidate: every slice equals the SRS prefix, byte for byte
    bool srs_confirm(const uint8_t* points, const uint32_t* counts, uint32_t nmsm);
    // host -> device through a ring of pinned chunks filled by several host
    // threads while earlier chunks are in flight (pageable memory moves at a
    // fraction of the link rate through the runtime's own staging)
    static constexpr uint32_t kStageSlots = 8;
    static constexpr uint64_t kStageChunk = 8ull << 20;
    int32_t stage_upload(void* dst, const uint8_t* src, uint64_t bytes);
    HostBuf stage_[kStageSlots];
    hipEvent_t stage_ev_[kStageSlots] = {};
    hipEvent_t ev_[kNumStages + 1] = {};
    hipEvent_t ev_tail_ = nullptr;  // after the device tail, before the copy-out
    // the part pipeline: the sort stream; timing events per part (sort,
    // accumulation and merge starts and ends; the reduction uses slot 0) and
    // the dependency events between the two streams
    hipStream_t stream_sort_ = nullptr;
    hipEvent_t ev_dec_ = nullptr;
    hipEvent_t ev_s0_[kMaxParts] = {}, ev_s1_[kMaxParts] = {}, ev_a0_[kMaxParts] = {},
               ev_a1_[kMaxParts] = {}, ev_m1_[kMaxParts] = {}, ev_r0_[kMaxParts] = {}, ev_r1_[kMaxParts] = {};
    hipEvent_t dep_fork_ = nullptr;
    hipEvent_t dep_sort_[kMaxParts] = {};
    uint32_t parts_req_ = 0;  // 0: default_parts
    // no-allocation path: this device's scratch array, a plan and host
    // staging whose capacity was reserved at creation
    uint8_t* scratch_ = nullptr;
    Plan na_plan_;
    std::vector<uint8_t> na_host_;
};

}  // namespace zk
