/*
 * zkmsm.h -- extensions of libzkmetal_amd.so beyond the zkMetal MSM ABI.
 * Same conventions as zkmetal.h (status codes, 64 B Montgomery affine
 * points, 32 B standard-form scalars, 96 B Jacobian results).
 *
 * None of these has a counterpart in the reference FFI; they expose the
 * MI355X-side structure the reference lacks: a device-resident base table
 * (the reference re-copies bases on every call, MSMEngine.swift:395-397),
 * batched MSMs over prefixes of it (KZGEngine.batchCommit,
 * Sources/zkMetal/KZG/KZGEngine.swift:151-179, is the natural caller), scalars
 * already in device memory, per-stage timing and a request server.
 */
#ifndef ZKMSM_H
#define ZKMSM_H

#include "zkmetal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device-resident base table (SRS) ---------------------------------- */

/* Upload n affine bases; the engine keeps [P_i, phi(P_i)] in HBM. */
ZkMetalStatus zkmsm_srs_upload(ZkMetalMSMEngine engine, const uint8_t* points, uint32_t n);
/* Synthetic table: P_i = k_i * G with k_i = zkmsm_gen_scalars(seed, n, 0)[i],
 * computed on the device (benchmark / test inputs). */
ZkMetalStatus zkmsm_srs_generate(ZkMetalMSMEngine engine, uint64_t seed, uint32_t n);
/* Copy the first n bases back (64 B each). */
/* Resident table from an SRS file image in the reference's layout
 * (Sources/zkMetal/KZG/SRSFileFormat.swift:57-160: "ptau", version 1,
 * sections {u32 type, u64 length}; section 1 header with curve id 1 = BN254;
 * section 2 G1 powers, 64 B per point, x and y 32-byte big-endian standard
 * form, (0,0) = identity).  The first max_points powers (0 = all) are
 * converted and checked on the device: a coordinate >= p or a point off the
 * curve fails the load with ZKMETAL_ERR_INVALID_INPUT (and leaves no table).
 * *n_loaded (nullable) receives the number of bases. */
ZkMetalStatus zkmsm_srs_load_ptau(ZkMetalMSMEngine engine, const uint8_t* data, uint64_t len,
                                  uint32_t max_points, uint32_t* n_loaded);
/* Same from a file path (memory-mapped). */
ZkMetalStatus zkmsm_srs_load_ptau_file(ZkMetalMSMEngine engine, const char* path,
                                       uint32_t max_points, uint32_t* n_loaded);
ZkMetalStatus zkmsm_srs_download(ZkMetalMSMEngine engine, uint8_t* points, uint32_t n);
uint32_t zkmsm_srs_size(ZkMetalMSMEngine engine);

/* MSM j uses bases [0, counts[j]) of the resident table and the next
 * counts[j] scalars of all_scalars.  counts[j] == 0 -> 96 zero bytes.
 * All MSMs run in one launch sequence. */
ZkMetalStatus zkmsm_msm_batch_resident(ZkMetalMSMEngine engine, const uint8_t* all_scalars,
                                       const uint32_t* counts, uint32_t n_msms,
                                       uint8_t* results);
/* Same, with the scalars already in device memory (from zkmsm_device_alloc). */
ZkMetalStatus zkmsm_msm_batch_resident_device(ZkMetalMSMEngine engine,
                                              const void* device_scalars,
                                              const uint32_t* counts, uint32_t n_msms,
                                              uint8_t* results);

/* The same as zkmsm_msm_batch_resident_device, allocating nothing: no
 * hipMalloc / hipFree / pinned-host allocation and no allocating library
 * call (no counterpart in the reference; for callers that cannot allocate:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  The call's device buffers come from a fixed array of 64 MiB per
 * device, in the library's device data; the host side uses capacity reserved
 * when the engine was created.  Limits: at most 65536 points over all MSMs,
 * at most 64 MSMs, and the call's buffers must fit the array (window widths
 * from the engine's tables; very small widths need more); a call beyond them
 * returns ZKMETAL_ERR_INVALID_INPUT without running.  Every engine on a
 * device shares that device's array: the library serialises such calls per
 * device (they do not overlap), and the array keeps the last call's
 * contents.  Results equal zkmsm_msm_batch_resident_device's bytes. */
ZkMetalStatus zkmsm_msm_batch_resident_device_noalloc(ZkMetalMSMEngine engine,
                                                      const void* device_scalars,
                                                      const uint32_t* counts, uint32_t n_msms,
                                                      uint8_t* results);

/* KZG commitments over the resident table (KZGEngine.commit / batchCommit,
 * Sources/zkMetal/KZG/KZGEngine.swift:131-179): polynomial p has counts[p]
 * coefficients, each an Fr element in Montgomery form (32 B, 8 x u32 LE --
 * the reference's Fr limbs), concatenated; commitment p = sum_j coeff_j *
 * SRS[j], 96 B X | Y | Z per polynomial as for zkmetal_bn254_msm.  The
 * Montgomery-to-standard conversion (the reference does it on the host,
 * bn254_fr_batch_to_limbs, Sources/NeonFieldOps/bn254_msm.c:1537-1552) runs
 * inside the decomposition kernel.  Polynomials of different lengths share
 * one call. */
ZkMetalStatus zkmsm_kzg_commit(ZkMetalMSMEngine engine, const uint8_t* coeffs_mont,
                               const uint32_t* counts, uint32_t n_polys, uint8_t* results);
/* Same, with the coefficients already in device memory (e.g. from a GPU NTT). */
ZkMetalStatus zkmsm_kzg_commit_device(ZkMetalMSMEngine engine, const void* device_coeffs_mont,
                                      const uint32_t* counts, uint32_t n_polys, uint8_t* results);

/* Grumpkin, BN254's cycle partner (y^2 = x^3 - 17 over BN254 Fr; the
 * reference's GrumpkinMSM.msm, Sources/zkMetal/MSM/GrumpkinMSMEngine.swift:246),
 * on the same engine.  points: n x 64 B affine (x, y in Fr Montgomery form,
 * 8 x u32 LE each; (0,0) = identity); scalars: n x 32 B standard form,
 * reduced mod the Grumpkin order (BN254's p) as the reference does
 * (GrumpkinMSMEngine.swift:178-196); result: Jacobian X | Y | Z in Fr
 * Montgomery form, normalised (Z = 1; the identity is (1, 1, 0)).  The batch
 * form takes concatenated inputs like zkmetal_bn254_msm_batch (a zero count
 * writes 96 zero bytes). */
ZkMetalStatus zkmsm_grumpkin_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                 const uint8_t* scalars, uint32_t n, uint8_t* X, uint8_t* Y,
                                 uint8_t* Z);
ZkMetalStatus zkmsm_grumpkin_msm_batch(ZkMetalMSMEngine engine, const uint8_t* all_points,
                                       const uint8_t* all_scalars, const uint32_t* counts,
                                       uint32_t n_msms, uint8_t* results);
/* The same with bases and scalars already in device memory (the Grumpkin
 * counterpart of zkmsm_msm_batch_device). */
ZkMetalStatus zkmsm_grumpkin_msm_batch_device(ZkMetalMSMEngine engine, const void* device_points,
                                              const void* device_scalars, const uint32_t* counts,
                                              uint32_t n_msms, uint8_t* results);
/* Synthetic Grumpkin bases: writes n points k_i * G (64 B each, Fr Montgomery,
 * (0,0) for k_i = 0) into device memory, k_i = zkmsm_gen_scalars(seed, n, 0)
 * and G = (1, sqrt(-16)) (GrumpkinCurve.swift:215-222).  For benchmarks and
 * tests, like zkmsm_srs_generate. */
ZkMetalStatus zkmsm_grumpkin_points_generate(ZkMetalMSMEngine engine, uint64_t seed, uint32_t n,
                                             void* device_points);

/* BN254 G2 (the reference's BN254G2MSM.msm,
 * Sources/zkMetal/MSM/BN254G2MSMEngine.swift:289) on the same engine.
 * points: n x 128 B affine, x.c0 | x.c1 | y.c0 | y.c1 (G2PointAffineGPU,
 * BN254G2MSMEngine.swift:11-27: Fp Montgomery form, 8 x u32 LE each; all zero
 * = identity); scalars: n x 32 B standard form, reduced mod r as the reference
 * does (BN254G2MSMEngine.swift:214-232); result: 192 B Jacobian
 * X | Y | Z, each c0 | c1 (G2PointProjectiveGPU, :29-43), normalised (Z = 1;
 * the identity is ((1,0), (1,0), (0,0)), g2Identity).  The batch forms take
 * concatenated inputs (results n_msms x 192 B, a zero count writes 192 zero
 * bytes); _device takes bases and scalars already in device memory. */
ZkMetalStatus zkmsm_bn254_g2_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                 const uint8_t* scalars, uint32_t n, uint8_t* result);
ZkMetalStatus zkmsm_bn254_g2_msm_batch(ZkMetalMSMEngine engine, const uint8_t* all_points,
                                       const uint8_t* all_scalars, const uint32_t* counts,
                                       uint32_t n_msms, uint8_t* results);
ZkMetalStatus zkmsm_bn254_g2_msm_batch_device(ZkMetalMSMEngine engine, const void* device_points,
                                              const void* device_scalars, const uint32_t* counts,
                                              uint32_t n_msms, uint8_t* results);
/* Synthetic G2 bases: n points k_i * G2 (128 B each) into device memory,
 * k_i = zkmsm_gen_scalars(seed, n, 0), G2 the reference's generator
 * (BN254Pairing.swift:389-396).  For benchmarks and tests. */
ZkMetalStatus zkmsm_bn254_g2_points_generate(ZkMetalMSMEngine engine, uint64_t seed, uint32_t n,
                                             void* device_points);

/* Host-point calls (zkmetal_bn254_msm, _batch, and the small-scalar
 * variants) whose point slices are each byte-for-byte a prefix of the
 * resident SRS run on the resident table: no base upload and no table build
 * (the batched launch mapping of zkMetal_ffi.swift:1241-1298).  The check is
 * exact (a 64-byte compare rejects other inputs at once; a match is confirmed
 * by comparing every slice with a host image of the SRS).  On by default;
 * zkmsm_srs_detect_hits counts the calls it served. */
ZkMetalStatus zkmsm_set_srs_detect(ZkMetalMSMEngine engine, int32_t on);
uint64_t zkmsm_srs_detect_hits(ZkMetalMSMEngine engine);

/* k MSMs sharing one base set (MultiMSM.swift:21-60 multiMSM): points n x 64 B,
 * scalar_sets k x n x 32 B (set j at offset j*n*32), results k x 96 B.  The
 * base table is built once; the k MSMs run as one batched launch. */
ZkMetalStatus zkmsm_multi_msm(ZkMetalMSMEngine engine, const uint8_t* points, uint32_t n,
                              const uint8_t* scalar_sets, uint32_t k, uint8_t* results);

/* Partitioned MSM (MultiMSM.swift:90-124 innerProductMSM): results[j] =
 * MSM(points[partitions[j] .. partitions[j+1]), same range of scalars);
 * partitions has n_parts + 1 entries, non-decreasing, first 0, last n.  An
 * empty partition writes 96 zero bytes (the identity). */
ZkMetalStatus zkmsm_inner_product_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                      const uint8_t* scalars, uint32_t n,
                                      const uint32_t* partitions, uint32_t n_parts,
                                      uint8_t* results);

/* Plain MSMs over bases and scalars that are both already in device memory
 * (the zkmetal_bn254_msm_batch contract without the host copies): the
 * [P, phi(P)] table is built inside the call. */
ZkMetalStatus zkmsm_msm_batch_device(ZkMetalMSMEngine engine, const void* device_points,
                                     const void* device_scalars, const uint32_t* counts,
                                     uint32_t n_msms, uint8_t* results);

/* ---- device memory for callers that keep scalars in HBM ----------------- */
ZkMetalStatus zkmsm_device_alloc(uint64_t bytes, void** device_ptr);
ZkMetalStatus zkmsm_device_free(void* device_ptr);
ZkMetalStatus zkmsm_copy_to_device(void* device_dst, const void* host_src, uint64_t bytes);
/* page-locked host memory: scalars submitted from it reach the device by DMA
 * without a staging copy */
ZkMetalStatus zkmsm_host_alloc(uint64_t bytes, void** host_ptr);
ZkMetalStatus zkmsm_host_free(void* host_ptr);
ZkMetalStatus zkmsm_copy_to_host(void* host_dst, const void* device_src, uint64_t bytes);

/* ---- tuning / introspection -------------------------------------------- */

/* Window width for subsequent calls on this engine; 0 restores the
 * per-size default.  Valid widths: 2..16. */
ZkMetalStatus zkmsm_set_window_bits(ZkMetalMSMEngine engine, uint32_t c);
/* Window width the default policy picks for an n-point MSM. */
uint32_t zkmsm_default_window_bits(uint64_t n);

/* Timing of the engine's last call.  out[0..5]: device ms of the stages
 * decompose, sort, accumulate, merge, bucket-reduce, copy-out; out[6]:
 * device total; out[7]: host tail ms; out[8]: wall ms of the call;
 * out[9]: window bits of the first MSM; out[10]: entries per accumulation
 * lane; out[11]: buckets per reduction group; out[12]: sorted entries;
 * out[13]: accumulation lanes; out[14]: entries with a nonzero digit (the
 * mixed additions performed); out[15]: device time of the window combine and
 * normalisation when the device ran them (0 when the host did, see
 * zkmsm_set_tail_mode); out[16]: pipeline parts of the call.  Returns the
 * number of values written (at most 17). */
int32_t zkmsm_last_timing(ZkMetalMSMEngine engine, double* out, int32_t max_values);

/* Per-size-class window widths: c_by_log2n[ceil(log2 n)] for 0 <= index < len
 * (<= 33 entries); 0 keeps the cost-model default for that class.  An engine
 * starts from the tables measured on MI355X (zkmetal_amd/csrc/tuned_gfx950.hpp);
 * these calls replace them.  The plain
 * table applies to calls with one MSM, the batch table to each MSM of a call
 * with several (falling back to the plain table where it holds 0).  Tables
 * come from the autotuning sweep in zkmetal_amd/tune.py; a request server
 * copies both tables from the engine it is created on. */
ZkMetalStatus zkmsm_set_window_table(ZkMetalMSMEngine engine, const uint32_t* c_by_log2n,
                                     uint32_t len);
ZkMetalStatus zkmsm_set_batch_window_table(ZkMetalMSMEngine engine, const uint32_t* c_by_log2n,
                                           uint32_t len);

/* Per-size-class launch schedule, indexed by ceil(log2 e) of a call's sorted
 * entries e (every (MSM, window) digit slot of the call, before zero digits
 * are dropped) for 0 <= index < len (<= 33 entries):
 *   parts_by_log2e[k]: pipeline parts the call's segments are split into
 *     (the sort of part p + 1 runs on a second stream beside the bucket
 *     accumulation of part p), 1..8;
 *   reduce_per_by_log2e[k]: buckets each bucket-reduction lane sums before
 *     the lanes of a row or column meet in a shuffle tree, a power of two
 *     in 4..64.
 * 0 keeps the built-in rule for that class; either array may be NULL (all
 * 0).  batch != 0 sets the table for calls of several MSMs, 0 for calls of
 * one.  An engine starts from the tables measured on MI355X
 * (zkmetal_amd/csrc/tuned_gfx950.hpp); they come from the autotuning sweep in
 * zkmetal_amd/tune.py, and a request server copies them from the engine it
 * is created on.  zkmsm_set_pipeline_parts, when nonzero, overrides the
 * parts.  Results do not depend on the schedule.  An out-of-range value
 * returns ZKMETAL_ERR_INVALID_INPUT and leaves the table unchanged.  (No
 * counterpart in the reference, whose threadgroup sizes come from
 * Sources/zkMetal/Tuning.swift:79-246.) */
ZkMetalStatus zkmsm_set_schedule_table(ZkMetalMSMEngine engine, int32_t batch,
                                       const uint32_t* parts_by_log2e,
                                       const uint32_t* reduce_per_by_log2e, uint32_t len);

/* Graph mode.  mode 1 (on): the device launch sequence of a call (all of
 * its MSMs: decomposition, sort, accumulation, merge, reduction, copy-out) is
 * captured once into a hipGraph and replayed by later calls of the same
 * shape (MSM sizes, window widths, table and scalar addresses); a call of a
 * new shape captures again, and the window combine runs on the device
 * (zkmsm_set_tail_mode auto).  mode 2 (auto, the default): the same from the
 * second call in a row of one shape (a one-off call runs eagerly and pays no
 * capture), with the tail where zkmsm_set_tail_mode puts it for eager calls.
 * mode 0 (off): every call eagerly.  Stage times in zkmsm_last_timing are 0
 * for a replay except the device total.  zkmsm_graph_captures counts the
 * captures.  Results are the same bytes in every mode. */
ZkMetalStatus zkmsm_set_graph_mode(ZkMetalMSMEngine engine, int32_t mode);

/* A large call runs as a pipeline over parts of its bucket segments: the
 * bucket sort of part p + 1 on a second stream while part p accumulates,
 * then one bucket reduction over all parts.  parts = 0 (default) picks 1, 2
 * or 3 parts by size; 1..8 forces that many (at most one per segment).
 * Results are the same bytes for any value. */
ZkMetalStatus zkmsm_set_pipeline_parts(ZkMetalMSMEngine engine, uint32_t parts);

/* Window-shifted resident tables.  mode 1 (auto, the default) or 2 (all):
 * an SRS set afterwards (zkmsm_srs_upload, zkmsm_srs_generate,
 * zkmsm_srs_load_ptau) also gets, for every 16-bit window w = 1..7, the
 * copies 2^(16 w) P_i and their endomorphism images, built on the device
 * (eight times the table's memory: 16 GiB for 2^23 bases).  An MSM over the
 * resident table can then accumulate all its windows into one set of 2^15
 * buckets: the bucket reduction shrinks eightfold and the window combine
 * needs no doublings, at the price of gathers spread over the larger table.
 * Auto takes it, in calls of several MSMs, for those whose cost estimate
 * favours it (the reduction saved against the padding added), for a single
 * MSM of at most 2^17 points, and for MSMs of >= 4096 points in calls whose
 * window combine runs on the device; all takes it for every MSM (measured
 * on MI355X: the 64-MSM batch 20.8 -> 18.0 ms, 32 witness-like MSMs of 2^18
 * points 13.1 -> 8.0 ms, one 2^16 MSM 0.61 -> 0.42 ms, one 2^22 MSM 6.05 ->
 * 6.23 ms, which auto runs plain).  mode 0 (off):
 * plain plans, and the next SRS is plain.  When the copies do not fit
 * (device memory, zkmsm_set_workspace_limit) the plain table is kept.
 * Results are the same bytes in every mode.  No counterpart in the
 * reference (its MSM reduces every window separately,
 * Sources/zkMetal/MSM/MSMEngine.swift:899-904). */
ZkMetalStatus zkmsm_set_window_shift(ZkMetalMSMEngine engine, int32_t mode);
/* 16 when the resident table holds window-shifted copies, else 0 */
uint32_t zkmsm_srs_window_shift(ZkMetalMSMEngine engine);

/* Where a call's window combine (Horner over the windows) and affine
 * normalisation run: 0 auto (the device in graph mode and for calls of 16 or
 * more MSMs, the host otherwise), 1 the host, 2 the device.  Results are the
 * same bytes either way.  BN254 G2 calls always use the host.  The device
 * form lets a graph replay end with final results; the host is faster for a
 * single MSM (the combine is a serial chain of ~127 doublings, which one
 * host core runs faster than one GPU lane). */
ZkMetalStatus zkmsm_set_tail_mode(ZkMetalMSMEngine engine, int32_t mode);

/* Cap on the device memory an engine's buffers may grow to (its resident SRS
 * included; 0 = no cap, the default), e.g. to share one GPU between
 * engines.  A call that would need more returns ZKMETAL_ERR_ALLOC_FAILED and
 * leaves the engine usable; zkmsm_workspace_bytes reports the bytes held.
 * No counterpart in the reference (its engine grows its buffers without a
 * bound, Sources/zkMetal/MSM/MSMEngine.swift:232-277). */
ZkMetalStatus zkmsm_set_workspace_limit(ZkMetalMSMEngine engine, uint64_t bytes);
uint64_t zkmsm_workspace_bytes(ZkMetalMSMEngine engine);
uint64_t zkmsm_graph_captures(ZkMetalMSMEngine engine);

/* ---- request server ------------------------------------------------------ */
/* Continuous batching of MSM requests over the resident table of an engine.
 * While a server exists, that engine's zkmsm_srs_upload / _generate /
 * _load_ptau[_file] return ZKMETAL_ERR_INVALID_INPUT (the table stays as the
 * server sees it), and zkmetal_msm_engine_destroy on it defers the release
 * until the last server is destroyed.  `slots` engines, each
 * with its own stream and workspace, take queued requests in arrival order
 * (at most max_batch per launch) while their estimated device footprint fits
 * the free HBM budget; batches on different slots overlap on the device.
 * Submission copies the scalars; all calls are thread-safe. */
typedef void* ZkmsmServer;
typedef struct {
    uint32_t slots;            /* 0 -> 2 */
    uint32_t max_batch;        /* 0 -> 64 */
    uint64_t hbm_budget_bytes; /* 0 -> 16 GiB */
} ZkmsmServerConfig;
typedef struct {
    uint64_t submitted, completed, failed, batches;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> result latency */
    double msms_per_s;                      /* completed / (last result - first submit) */
    double mean_batch;
    uint64_t hbm_peak_bytes, hbm_in_use_bytes; /* admitted footprint */
} ZkmsmServerStats;
ZkMetalStatus zkmsm_server_create(ZkMetalMSMEngine srs_engine, const ZkmsmServerConfig* cfg,
                                  ZkmsmServer* out);
/* n in [1, SRS size]; *ticket identifies the request */
ZkMetalStatus zkmsm_server_submit(ZkmsmServer server, const uint8_t* scalars, uint32_t n,
                                  uint64_t* ticket);
/* same without the copy: the caller keeps `scalars` valid and unchanged until
 * the request is done (wait returns or poll reports done) */
ZkMetalStatus zkmsm_server_submit_ref(ZkmsmServer server, const uint8_t* scalars, uint32_t n,
                                      uint64_t* ticket);
/* block until the request is done; result: 96 B X|Y|Z; releases the ticket */
ZkMetalStatus zkmsm_server_wait(ZkmsmServer server, uint64_t ticket, uint8_t* result,
                                double* latency_ms);
/* non-blocking: *done = 0 while pending; on done, result written, ticket released */
ZkMetalStatus zkmsm_server_poll(ZkmsmServer server, uint64_t ticket, uint8_t* result,
                                int32_t* done);
ZkMetalStatus zkmsm_server_stats(ZkmsmServer server, ZkmsmServerStats* out);
void zkmsm_server_destroy(ZkmsmServer server);
/* bytes of the HBM budget a request of n points is charged */
uint64_t zkmsm_server_footprint(uint32_t n);

/* ---- seeded synthetic scalars (host) ------------------------------------ */
/* dist 0: uniform in Fr; 1: 40% 0, 30% < 2^16, 10% in {1, r-1}, 20% uniform;
 * 2: uniform 256-bit.  out: n x 32 B. */
void zkmsm_gen_scalars(uint64_t seed, uint64_t n, int32_t dist, uint8_t* out);

/* ---- DIAGNOSTICS: not part of the MSM API ------------------------------- */
/* The zkmsm_test_* and zkmsm_probe_* entry points below exist for the test
 * suite and for measurement (primitive parity tests on host buffers, the
 * roofline peak and the HBM counter calibration).  A caller of the MSM path
 * never needs them; their signatures may change between versions. */

/* ---- primitive probes for GPU unit tests (host buffers) ------------------ */
/* op 0 mul, 1 add, 2 sub, 3 inverse(a), 4 negate(a), 5 canonicalise(a);
 * a, b, out: n x 32 B Montgomery field elements. */
ZkMetalStatus zkmsm_test_fp(ZkMetalMSMEngine engine, int32_t op, const uint8_t* a,
                            const uint8_t* b, uint8_t* out, uint32_t n);
/* op 0: p + q, 1: 2p, 2: p + affine(q.X, q.Y); XYZZ points of 128 B. */
ZkMetalStatus zkmsm_test_xyzz(ZkMetalMSMEngine engine, int32_t op, const uint8_t* p,
                              const uint8_t* q, uint8_t* out, uint32_t n);
/* BN254 G2 primitives on 256-byte XYZZ points (X, Y, ZZ, ZZZ c0 parts, then
 * the c1 parts; Fp Montgomery form).  op 0: p + q, 1: 2p, 2: p + affine
 * (q.X, q.Y) (all zero = identity), 3: the bucket-loop addition p + (q.X, q.Y)
 * (p the identity = a run start); field ops on X, result in X, the rest zero:
 * 4: p.X * q.X, 5: p.X^2, 6: 1 / p.X. */
ZkMetalStatus zkmsm_test_g2(ZkMetalMSMEngine engine, int32_t op, const uint8_t* p,
                            const uint8_t* q, uint8_t* out, uint32_t n);
/* GLV split of n scalars: out n x 48 B = k1 (16 B), k2 (16 B), neg1, neg2 (u32). */
ZkMetalStatus zkmsm_test_glv(ZkMetalMSMEngine engine, const uint8_t* scalars, uint8_t* out,
                             uint32_t n);

/* Measured device rates: out6[0] = v_mad_u64_u32 (32x32->64 multiply-add)
 * per second at full occupancy -- the INT32-multiply roofline peak;
 * out6[1] = FIPS accumulate steps (multiply-add + carry add) per second;
 * out6[2] = 256-bit Montgomery multiplies per second (8 x 32-bit limbs);
 * out6[3] = ns per dependent multiply in a single lane (one wave per SIMD);
 * out6[4], out6[5] = the same two for the 9 x 29-bit limb multiply. */
ZkMetalStatus zkmsm_probe_rates(ZkMetalMSMEngine engine, double* out6);
/* One launch of a plain device copy of `bytes` bytes (read + write), timed
 * with HIP events; the calibration point for HBM byte counters. */
ZkMetalStatus zkmsm_probe_copy(ZkMetalMSMEngine engine, uint64_t bytes, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* ZKMSM_H */
