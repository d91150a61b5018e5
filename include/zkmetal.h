/*
 * zkmetal.h -- BN254 G1 MSM C ABI of libzkmetal_amd.so (MI355X / gfx950).
 *
 * Drop-in for the MSM subset of zkMetal's C FFI
 * (Sources/zkMetal-ffi/include/zkmetal.h); names, signatures, status codes,
 * byte layouts and ownership rules are the same, so callers written against
 * that header (Go cgo bindings/go/zkmetal/gpu_msm.go:29-76, Rust
 * bindings/rust/src/msm.rs:78-242, C bindings/c/) link against this library
 * unchanged.  Engines for other curves, NTT, hashes, FRI and pairings are
 * not part of this library.
 *
 * Layouts (reference zkmetal.h:1-11):
 *   field element  32 B, 8 x u32 little-endian, Montgomery form
 *   affine point   64 B (x, y); (0,0) is the point at infinity
 *   scalar         32 B, 8 x u32 little-endian, STANDARD form, any 256-bit
 *                  value (reduced mod r internally)
 *   result         Jacobian X, Y, Z (32 B each, Montgomery).  This library
 *                  always returns the normalised representative: Z = 1 (in
 *                  Montgomery form) for a finite point, (1, 1, 0) for the
 *                  identity.
 * Buffers are caller-owned and not retained after a call returns.
 */
#ifndef ZKMETAL_H
#define ZKMETAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (reference zkmetal.h:28-34) */
typedef int32_t ZkMetalStatus;
#define ZKMETAL_SUCCESS 0
#define ZKMETAL_ERR_NO_GPU -1
#define ZKMETAL_ERR_INVALID_INPUT -2
#define ZKMETAL_ERR_GPU_ERROR -3
#define ZKMETAL_ERR_ALLOC_FAILED -4

/* Opaque engine handle (reference zkmetal.h:40).  One engine owns a HIP
 * stream and a grow-only device workspace on the current device; calls on
 * one engine are serialised internally, distinct engines run concurrently. */
typedef void* ZkMetalMSMEngine;

/* reference zkmetal.h:57; ZKMETAL_ERR_NO_GPU when no HIP device exists */
ZkMetalStatus zkmetal_msm_engine_create(ZkMetalMSMEngine* out);
/* reference zkmetal.h:60; NULL is ignored */
void zkmetal_msm_engine_destroy(ZkMetalMSMEngine engine);

/* result = sum scalars[i] * points[i]  (reference zkmetal.h:129-137).
 * n_points == 0 -> ZKMETAL_ERR_INVALID_INPUT (zkMetal_ffi.swift:349). */
ZkMetalStatus zkmetal_bn254_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                const uint8_t* scalars, uint32_t n_points, uint8_t* result_x,
                                uint8_t* result_y, uint8_t* result_z);

/* Same, on a lazily created process-wide engine (reference zkmetal.h:140-147). */
ZkMetalStatus zkmetal_bn254_msm_auto(const uint8_t* points, const uint8_t* scalars,
                                     uint32_t n_points, uint8_t* result_x, uint8_t* result_y,
                                     uint8_t* result_z);

/* Small-scalar variants: scalars are n_points little-endian u8 / u16 / u32
 * values, zero-extended (reference zkmetal.h:446-484, zkMetal_ffi.swift:1152-1206). */
ZkMetalStatus zkmetal_bn254_msm_u8(ZkMetalMSMEngine engine, const uint8_t* points,
                                   const uint8_t* scalars, uint32_t n_points, uint8_t* result_x,
                                   uint8_t* result_y, uint8_t* result_z);
ZkMetalStatus zkmetal_bn254_msm_u16(ZkMetalMSMEngine engine, const uint8_t* points,
                                    const uint8_t* scalars, uint32_t n_points, uint8_t* result_x,
                                    uint8_t* result_y, uint8_t* result_z);
ZkMetalStatus zkmetal_bn254_msm_u32(ZkMetalMSMEngine engine, const uint8_t* points,
                                    const uint8_t* scalars, uint32_t n_points, uint8_t* result_x,
                                    uint8_t* result_y, uint8_t* result_z);
ZkMetalStatus zkmetal_bn254_msm_u8_auto(const uint8_t* points, const uint8_t* scalars,
                                        uint32_t n_points, uint8_t* result_x, uint8_t* result_y,
                                        uint8_t* result_z);
ZkMetalStatus zkmetal_bn254_msm_u16_auto(const uint8_t* points, const uint8_t* scalars,
                                         uint32_t n_points, uint8_t* result_x, uint8_t* result_y,
                                         uint8_t* result_z);
ZkMetalStatus zkmetal_bn254_msm_u32_auto(const uint8_t* points, const uint8_t* scalars,
                                         uint32_t n_points, uint8_t* result_x, uint8_t* result_y,
                                         uint8_t* result_z);

/* Batch of independent MSMs over concatenated inputs (reference
 * zkmetal.h:499-515).  results: n_msms x 96 B (X|Y|Z).  counts[i] == 0
 * writes 96 zero bytes for that entry (zkMetal_ffi.swift:1260-1264);
 * n_msms == 0 -> ZKMETAL_ERR_INVALID_INPUT.  Unlike the reference's
 * sequential loop, all MSMs of the call run in one launch sequence. */
ZkMetalStatus zkmetal_bn254_msm_batch(ZkMetalMSMEngine engine, const uint8_t* all_points,
                                      const uint8_t* all_scalars, const uint32_t* counts,
                                      uint32_t n_msms, uint8_t* results);
ZkMetalStatus zkmetal_bn254_msm_batch_auto(const uint8_t* all_points, const uint8_t* all_scalars,
                                           const uint32_t* counts, uint32_t n_msms,
                                           uint8_t* results);

/* Utilities (reference zkmetal.h:524-530).  There are no shader files on
 * HIP (the kernels are compiled into the library): set_shader_dir only
 * records the path; gpu_available reports whether a HIP device exists. */
void zkmetal_set_shader_dir(const char* path);
int32_t zkmetal_gpu_available(void);
const char* zkmetal_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ZKMETAL_H */
