// extern "C" boundary of libzkmetal_amd.so: the zkMetal MSM ABI
// (include/zkmetal.h) and this library's extensions (include/zkmsm.h).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/zkmsm.h"
#include "engine.hpp"
#include "gen.hpp"
#include "server.hpp"

using zk::Engine;

namespace {

std::mutex g_auto_mu;
Engine* g_auto = nullptr;
std::mutex g_dir_mu;
std::string g_shader_dir;

// lazily created process-wide engine (zkMetal_ffi.swift:23-52)
int32_t auto_engine(Engine** e) {
    std::lock_guard<std::mutex> lk(g_auto_mu);
    if (!g_auto) {
        int32_t rc = Engine::create(&g_auto);
        if (rc != zk::kOk) return rc;
    }
    *e = g_auto;
    return zk::kOk;
}

inline Engine* as_engine(ZkMetalMSMEngine h) { return static_cast<Engine*>(h); }

int32_t split_out(const uint8_t* r96, uint8_t* x, uint8_t* y, uint8_t* z) {
    memcpy(x, r96, 32);
    memcpy(y, r96 + 32, 32);
    memcpy(z, r96 + 64, 32);
    return zk::kOk;
}

int32_t msm_impl(Engine* e, const uint8_t* points, const uint8_t* scalars, uint32_t n,
                 uint8_t* x, uint8_t* y, uint8_t* z) {
    if (!e) return zk::kInvalidInput;
    if (n == 0 || !points || !scalars || !x || !y || !z) return zk::kInvalidInput;
    uint8_t r[96];
    int32_t rc;
    {
        std::lock_guard<std::mutex> lk(e->mutex());
        rc = e->msm_host(points, scalars, n, r);
    }
    if (rc != zk::kOk) return rc;
    return split_out(r, x, y, z);
}

// zero-extend 1/2/4-byte little-endian scalars (zkMetal_ffi.swift:1176-1196)
int32_t msm_small_impl(Engine* e, const uint8_t* points, const uint8_t* scalars, uint32_t n,
                       int width, uint8_t* x, uint8_t* y, uint8_t* z) {
    if (n == 0 || !scalars) return zk::kInvalidInput;
    std::vector<uint8_t> full((size_t)n * 32, 0);
    for (uint32_t i = 0; i < n; i++) memcpy(&full[(size_t)i * 32], scalars + (size_t)i * width, width);
    return msm_impl(e, points, full.data(), n, x, y, z);
}

int32_t batch_impl(Engine* e, const uint8_t* pts, const uint8_t* sc, const uint32_t* counts,
                   uint32_t n_msms, uint8_t* results) {
    if (!e) return zk::kInvalidInput;
    if (n_msms == 0) return zk::kInvalidInput;  // zkMetal_ffi.swift:1250
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_batch_host(pts, sc, counts, n_msms, results);
}

}  // namespace

extern "C" {

ZkMetalStatus zkmetal_msm_engine_create(ZkMetalMSMEngine* out) {
    if (!out) return ZKMETAL_ERR_INVALID_INPUT;
    Engine* e = nullptr;
    int32_t rc = Engine::create(&e);
    *out = e;
    return rc;
}

// a request server created on this engine keeps it alive until the server is
// destroyed (Engine::release)
void zkmetal_msm_engine_destroy(ZkMetalMSMEngine engine) { Engine::release(as_engine(engine)); }

ZkMetalStatus zkmetal_bn254_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                const uint8_t* scalars, uint32_t n_points, uint8_t* result_x,
                                uint8_t* result_y, uint8_t* result_z) {
    return msm_impl(as_engine(engine), points, scalars, n_points, result_x, result_y, result_z);
}

ZkMetalStatus zkmetal_bn254_msm_auto(const uint8_t* points, const uint8_t* scalars,
                                     uint32_t n_points, uint8_t* result_x, uint8_t* result_y,
                                     uint8_t* result_z) {
    Engine* e = nullptr;
    int32_t rc = auto_engine(&e);
    if (rc != zk::kOk) return rc;
    return msm_impl(e, points, scalars, n_points, result_x, result_y, result_z);
}

#define ZK_SMALL(name, width)                                                                   \
    ZkMetalStatus name(ZkMetalMSMEngine engine, const uint8_t* points, const uint8_t* scalars, \
                       uint32_t n, uint8_t* x, uint8_t* y, uint8_t* z) {                        \
        return msm_small_impl(as_engine(engine), points, scalars, n, width, x, y, z);          \
    }                                                                                           \
    ZkMetalStatus name##_auto(const uint8_t* points, const uint8_t* scalars, uint32_t n,      \
                              uint8_t* x, uint8_t* y, uint8_t* z) {                             \
        Engine* e = nullptr;                                                                    \
        int32_t rc = auto_engine(&e);                                                           \
        if (rc != zk::kOk) return rc;                                                           \
        return msm_small_impl(e, points, scalars, n, width, x, y, z);                          \
    }
ZK_SMALL(zkmetal_bn254_msm_u8, 1)
ZK_SMALL(zkmetal_bn254_msm_u16, 2)
ZK_SMALL(zkmetal_bn254_msm_u32, 4)
#undef ZK_SMALL

ZkMetalStatus zkmetal_bn254_msm_batch(ZkMetalMSMEngine engine, const uint8_t* all_points,
                                      const uint8_t* all_scalars, const uint32_t* counts,
                                      uint32_t n_msms, uint8_t* results) {
    return batch_impl(as_engine(engine), all_points, all_scalars, counts, n_msms, results);
}

ZkMetalStatus zkmetal_bn254_msm_batch_auto(const uint8_t* all_points, const uint8_t* all_scalars,
                                           const uint32_t* counts, uint32_t n_msms,
                                           uint8_t* results) {
    Engine* e = nullptr;
    int32_t rc = auto_engine(&e);
    if (rc != zk::kOk) return rc;
    return batch_impl(e, all_points, all_scalars, counts, n_msms, results);
}

ZkMetalStatus zkmsm_grumpkin_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                 const uint8_t* scalars, uint32_t n, uint8_t* x, uint8_t* y,
                                 uint8_t* z) {
    Engine* e = as_engine(engine);
    if (!e || n == 0 || !points || !scalars || !x || !y || !z) return ZKMETAL_ERR_INVALID_INPUT;
    uint8_t r[96];
    int32_t rc;
    {
        std::lock_guard<std::mutex> lk(e->mutex());
        rc = e->msm_batch_host(points, scalars, &n, 1, r, zk::CurveId::Grumpkin);
    }
    if (rc != zk::kOk) return rc;
    return split_out(r, x, y, z);
}

ZkMetalStatus zkmsm_grumpkin_msm_batch(ZkMetalMSMEngine engine, const uint8_t* all_points,
                                       const uint8_t* all_scalars, const uint32_t* counts,
                                       uint32_t n_msms, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e || n_msms == 0) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_batch_host(all_points, all_scalars, counts, n_msms, results,
                             zk::CurveId::Grumpkin);
}

ZkMetalStatus zkmsm_grumpkin_msm_batch_device(ZkMetalMSMEngine engine, const void* device_points,
                                              const void* device_scalars, const uint32_t* counts,
                                              uint32_t n_msms, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_device(device_points, device_scalars, counts, n_msms, results,
                         zk::CurveId::Grumpkin);
}

ZkMetalStatus zkmsm_grumpkin_points_generate(ZkMetalMSMEngine engine, uint64_t seed, uint32_t n,
                                             void* device_points) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->grumpkin_points_generate(seed, n, device_points);
}

ZkMetalStatus zkmsm_bn254_g2_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                 const uint8_t* scalars, uint32_t n, uint8_t* result) {
    Engine* e = as_engine(engine);
    if (!e || n == 0 || !points || !scalars || !result) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_batch_host(points, scalars, &n, 1, result, zk::CurveId::Bn254G2);
}

ZkMetalStatus zkmsm_bn254_g2_msm_batch(ZkMetalMSMEngine engine, const uint8_t* all_points,
                                       const uint8_t* all_scalars, const uint32_t* counts,
                                       uint32_t n_msms, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e || n_msms == 0) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_batch_host(all_points, all_scalars, counts, n_msms, results, zk::CurveId::Bn254G2);
}

ZkMetalStatus zkmsm_bn254_g2_msm_batch_device(ZkMetalMSMEngine engine, const void* device_points,
                                              const void* device_scalars, const uint32_t* counts,
                                              uint32_t n_msms, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_device(device_points, device_scalars, counts, n_msms, results, zk::CurveId::Bn254G2);
}

ZkMetalStatus zkmsm_bn254_g2_points_generate(ZkMetalMSMEngine engine, uint64_t seed, uint32_t n,
                                             void* device_points) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->g2_points_generate(seed, n, device_points);
}

ZkMetalStatus zkmsm_set_srs_detect(ZkMetalMSMEngine engine, int32_t on) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    e->set_srs_detect(on != 0);
    return ZKMETAL_SUCCESS;
}

uint64_t zkmsm_srs_detect_hits(ZkMetalMSMEngine engine) {
    Engine* e = as_engine(engine);
    if (!e) return 0;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->srs_detect_hits();
}

ZkMetalStatus zkmsm_multi_msm(ZkMetalMSMEngine engine, const uint8_t* points, uint32_t n,
                              const uint8_t* scalar_sets, uint32_t k, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->multi_msm_host(points, n, scalar_sets, k, results);
}

ZkMetalStatus zkmsm_inner_product_msm(ZkMetalMSMEngine engine, const uint8_t* points,
                                      const uint8_t* scalars, uint32_t n,
                                      const uint32_t* partitions, uint32_t n_parts,
                                      uint8_t* results) {
    // MultiMSM.swift:90-124: boundaries sorted, first 0, last n; an empty
    // partition gives the identity (96 zero bytes, as a zero count does in
    // the batch call)
    Engine* e = as_engine(engine);
    if (!e || !partitions || n_parts == 0 || partitions[0] != 0 || partitions[n_parts] != n)
        return ZKMETAL_ERR_INVALID_INPUT;
    std::vector<uint32_t> counts(n_parts);
    for (uint32_t j = 0; j < n_parts; j++) {
        if (partitions[j + 1] < partitions[j]) return ZKMETAL_ERR_INVALID_INPUT;
        counts[j] = partitions[j + 1] - partitions[j];
    }
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_batch_host(points, scalars, counts.data(), n_parts, results);
}

void zkmetal_set_shader_dir(const char* path) {
    std::lock_guard<std::mutex> lk(g_dir_mu);
    g_shader_dir = path ? path : "";
}

int32_t zkmetal_gpu_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count > 0 ? 1 : 0;
}

const char* zkmetal_version(void) { return "0.2.0-mi355x"; }

// ---------------------------------------------------------------- extensions

ZkMetalStatus zkmsm_srs_upload(ZkMetalMSMEngine engine, const uint8_t* points, uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->srs_upload(points, n);
}

ZkMetalStatus zkmsm_srs_generate(ZkMetalMSMEngine engine, uint64_t seed, uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->srs_generate(seed, n);
}

ZkMetalStatus zkmsm_srs_download(ZkMetalMSMEngine engine, uint8_t* points, uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->srs_download(points, n);
}

uint32_t zkmsm_srs_size(ZkMetalMSMEngine engine) {
    Engine* e = as_engine(engine);
    return e ? e->srs_size() : 0;
}

ZkMetalStatus zkmsm_msm_batch_resident(ZkMetalMSMEngine engine, const uint8_t* all_scalars,
                                       const uint32_t* counts, uint32_t n_msms,
                                       uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_resident(all_scalars, false, counts, n_msms, results);
}

ZkMetalStatus zkmsm_msm_batch_resident_device(ZkMetalMSMEngine engine,
                                              const void* device_scalars,
                                              const uint32_t* counts, uint32_t n_msms,
                                              uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_resident(device_scalars, true, counts, n_msms, results);
}

ZkMetalStatus zkmsm_msm_batch_resident_device_noalloc(ZkMetalMSMEngine engine,
                                                      const void* device_scalars,
                                                      const uint32_t* counts, uint32_t n_msms,
                                                      uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_resident_noalloc(device_scalars, counts, n_msms, results);
}

ZkMetalStatus zkmsm_srs_load_ptau(ZkMetalMSMEngine engine, const uint8_t* data, uint64_t len,
                                  uint32_t max_points, uint32_t* n_loaded) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->srs_load_ptau(data, len, max_points, n_loaded);
}

ZkMetalStatus zkmsm_srs_load_ptau_file(ZkMetalMSMEngine engine, const char* path,
                                       uint32_t max_points, uint32_t* n_loaded) {
    Engine* e = as_engine(engine);
    if (!e || !path) return ZKMETAL_ERR_INVALID_INPUT;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return ZKMETAL_ERR_INVALID_INPUT;
    struct stat sb;
    if (fstat(fd, &sb) != 0 || sb.st_size < 12) {
        close(fd);
        return ZKMETAL_ERR_INVALID_INPUT;
    }
    void* map = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return ZKMETAL_ERR_INVALID_INPUT;
    ZkMetalStatus rc;
    {
        std::lock_guard<std::mutex> lk(e->mutex());
        rc = e->srs_load_ptau((const uint8_t*)map, (uint64_t)sb.st_size, max_points, n_loaded);
    }
    munmap(map, (size_t)sb.st_size);
    return rc;
}

ZkMetalStatus zkmsm_kzg_commit(ZkMetalMSMEngine engine, const uint8_t* coeffs_mont,
                               const uint32_t* counts, uint32_t n_polys, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->kzg_commit(coeffs_mont, false, counts, n_polys, results);
}

ZkMetalStatus zkmsm_kzg_commit_device(ZkMetalMSMEngine engine, const void* device_coeffs_mont,
                                      const uint32_t* counts, uint32_t n_polys, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->kzg_commit(device_coeffs_mont, true, counts, n_polys, results);
}

ZkMetalStatus zkmsm_msm_batch_device(ZkMetalMSMEngine engine, const void* device_points,
                                     const void* device_scalars, const uint32_t* counts,
                                     uint32_t n_msms, uint8_t* results) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->msm_device(device_points, device_scalars, counts, n_msms, results);
}

ZkMetalStatus zkmsm_device_alloc(uint64_t bytes, void** device_ptr) {
    if (!device_ptr || bytes == 0) return ZKMETAL_ERR_INVALID_INPUT;
    if (hipMalloc(device_ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *device_ptr = nullptr;
        return ZKMETAL_ERR_ALLOC_FAILED;
    }
    return ZKMETAL_SUCCESS;
}

ZkMetalStatus zkmsm_device_free(void* device_ptr) {
    if (!device_ptr) return ZKMETAL_SUCCESS;
    return hipFree(device_ptr) == hipSuccess ? ZKMETAL_SUCCESS : ZKMETAL_ERR_GPU_ERROR;
}

ZkMetalStatus zkmsm_host_alloc(uint64_t bytes, void** host_ptr) {
    if (!host_ptr || bytes == 0) return ZKMETAL_ERR_INVALID_INPUT;
    if (hipHostMalloc(host_ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *host_ptr = nullptr;
        return ZKMETAL_ERR_ALLOC_FAILED;
    }
    return ZKMETAL_SUCCESS;
}

ZkMetalStatus zkmsm_host_free(void* host_ptr) {
    if (!host_ptr) return ZKMETAL_SUCCESS;
    return hipHostFree(host_ptr) == hipSuccess ? ZKMETAL_SUCCESS : ZKMETAL_ERR_GPU_ERROR;
}

ZkMetalStatus zkmsm_copy_to_device(void* dst, const void* src, uint64_t bytes) {
    if (!dst || !src) return ZKMETAL_ERR_INVALID_INPUT;
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? ZKMETAL_SUCCESS
                                                                           : ZKMETAL_ERR_GPU_ERROR;
}

ZkMetalStatus zkmsm_copy_to_host(void* dst, const void* src, uint64_t bytes) {
    if (!dst || !src) return ZKMETAL_ERR_INVALID_INPUT;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? ZKMETAL_SUCCESS
                                                                           : ZKMETAL_ERR_GPU_ERROR;
}

ZkMetalStatus zkmsm_set_window_bits(ZkMetalMSMEngine engine, uint32_t c) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    if (c != 0 && (c < 2 || c > 16)) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    e->set_window_bits(c);
    return ZKMETAL_SUCCESS;
}

uint32_t zkmsm_default_window_bits(uint64_t n) { return zk::default_window_bits(n); }

int32_t zkmsm_last_timing(ZkMetalMSMEngine engine, double* out, int32_t max_values) {
    Engine* e = as_engine(engine);
    if (!e || !out) return 0;
    std::lock_guard<std::mutex> lk(e->mutex());
    const zk::Timing& t = e->last_timing();
    double v[17];
    for (int s = 0; s < zk::kNumStages; s++) v[s] = t.stage_ms[s];
    v[6] = t.device_ms;
    v[7] = t.host_tail_ms;
    v[8] = t.wall_ms;
    v[9] = t.window_bits_first;
    v[10] = t.T;
    v[11] = t.G;
    v[12] = (double)t.entries;
    v[13] = (double)t.lanes;
    v[14] = (double)t.valid_entries;
    v[15] = t.device_tail_ms;
    v[16] = t.parts;
    int32_t k = max_values < 17 ? max_values : 17;
    for (int32_t i = 0; i < k; i++) out[i] = v[i];
    return k;
}

void zkmsm_gen_scalars(uint64_t seed, uint64_t n, int32_t dist, uint8_t* out) {
    if (!out) return;
    zkh::gen_scalars(seed, n, dist, out);
}

ZkMetalStatus zkmsm_test_fp(ZkMetalMSMEngine engine, int32_t op, const uint8_t* a,
                            const uint8_t* b, uint8_t* out, uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->test_fp(op, a, b, out, n);
}

ZkMetalStatus zkmsm_test_g2(ZkMetalMSMEngine engine, int32_t op, const uint8_t* p,
                            const uint8_t* q, uint8_t* out, uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->test_g2(op, p, q, out, n);
}

ZkMetalStatus zkmsm_test_xyzz(ZkMetalMSMEngine engine, int32_t op, const uint8_t* p,
                              const uint8_t* q, uint8_t* out, uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->test_xyzz(op, p, q, out, n);
}

ZkMetalStatus zkmsm_test_glv(ZkMetalMSMEngine engine, const uint8_t* scalars, uint8_t* out,
                             uint32_t n) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->test_glv(scalars, out, n);
}

ZkMetalStatus zkmsm_probe_rates(ZkMetalMSMEngine engine, double* out6) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->probe_rates(out6);
}

ZkMetalStatus zkmsm_probe_copy(ZkMetalMSMEngine engine, uint64_t bytes, double* ms) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->probe_copy(bytes, ms);
}

ZkMetalStatus zkmsm_set_window_table(ZkMetalMSMEngine engine, const uint32_t* c_by_log2n,
                                     uint32_t len) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->set_window_table(c_by_log2n, len, false);
}

ZkMetalStatus zkmsm_set_pipeline_parts(ZkMetalMSMEngine engine, uint32_t parts) {
    Engine* e = as_engine(engine);
    if (!e || parts > zk::kMaxParts) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    e->set_parts(parts);
    return ZKMETAL_SUCCESS;
}

ZkMetalStatus zkmsm_set_window_shift(ZkMetalMSMEngine engine, int32_t mode) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->set_window_shift(mode);
}

uint32_t zkmsm_srs_window_shift(ZkMetalMSMEngine engine) {
    Engine* e = as_engine(engine);
    if (!e) return 0;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->srs_window_shift();
}

ZkMetalStatus zkmsm_set_tail_mode(ZkMetalMSMEngine engine, int32_t mode) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->set_tail_mode(mode);
}

ZkMetalStatus zkmsm_set_workspace_limit(ZkMetalMSMEngine engine, uint64_t bytes) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    e->set_workspace_limit(bytes);
    return ZKMETAL_SUCCESS;
}

uint64_t zkmsm_workspace_bytes(ZkMetalMSMEngine engine) {
    Engine* e = as_engine(engine);
    if (!e) return 0;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->workspace_bytes();
}

ZkMetalStatus zkmsm_set_graph_mode(ZkMetalMSMEngine engine, int32_t mode) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->set_graph_mode(mode);
}

uint64_t zkmsm_graph_captures(ZkMetalMSMEngine engine) {
    Engine* e = as_engine(engine);
    if (!e) return 0;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->graph_captures();
}

ZkMetalStatus zkmsm_set_batch_window_table(ZkMetalMSMEngine engine, const uint32_t* c_by_log2n,
                                           uint32_t len) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->set_window_table(c_by_log2n, len, true);
}

ZkMetalStatus zkmsm_set_schedule_table(ZkMetalMSMEngine engine, int32_t batch, const uint32_t* parts_by_log2e,
                                       const uint32_t* reduce_per_by_log2e, uint32_t len) {
    Engine* e = as_engine(engine);
    if (!e) return ZKMETAL_ERR_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(e->mutex());
    return e->set_schedule_table(parts_by_log2e, reduce_per_by_log2e, len, batch != 0);
}

// ---------------------------------------------------------------- server

ZkMetalStatus zkmsm_server_create(ZkMetalMSMEngine srs_engine, const ZkmsmServerConfig* cfg,
                                  ZkmsmServer* out) {
    Engine* e = as_engine(srs_engine);
    if (!e || !out) return ZKMETAL_ERR_INVALID_INPUT;
    zk::ServerConfig sc;
    if (cfg) {
        if (cfg->slots) sc.slots = cfg->slots;
        if (cfg->max_batch) sc.max_batch = cfg->max_batch;
        if (cfg->hbm_budget_bytes) sc.hbm_budget = cfg->hbm_budget_bytes;
    }
    std::lock_guard<std::mutex> lk(e->mutex());
    zk::Server* s = nullptr;
    int32_t rc = zk::Server::create(e, sc, &s);  // slots take e's window tables
    *out = s;
    return rc;
}

ZkMetalStatus zkmsm_server_submit(ZkmsmServer server, const uint8_t* scalars, uint32_t n,
                                  uint64_t* ticket) {
    zk::Server* s = static_cast<zk::Server*>(server);
    if (!s) return ZKMETAL_ERR_INVALID_INPUT;
    return s->submit(scalars, n, ticket, true);
}

ZkMetalStatus zkmsm_server_submit_ref(ZkmsmServer server, const uint8_t* scalars, uint32_t n,
                                      uint64_t* ticket) {
    zk::Server* s = static_cast<zk::Server*>(server);
    if (!s) return ZKMETAL_ERR_INVALID_INPUT;
    return s->submit(scalars, n, ticket, false);
}

ZkMetalStatus zkmsm_server_wait(ZkmsmServer server, uint64_t ticket, uint8_t* result,
                                double* latency_ms) {
    zk::Server* s = static_cast<zk::Server*>(server);
    if (!s) return ZKMETAL_ERR_INVALID_INPUT;
    return s->wait(ticket, result, latency_ms);
}

ZkMetalStatus zkmsm_server_poll(ZkmsmServer server, uint64_t ticket, uint8_t* result,
                                int32_t* done) {
    zk::Server* s = static_cast<zk::Server*>(server);
    if (!s) return ZKMETAL_ERR_INVALID_INPUT;
    return s->poll(ticket, result, done);
}

ZkMetalStatus zkmsm_server_stats(ZkmsmServer server, ZkmsmServerStats* out) {
    zk::Server* s = static_cast<zk::Server*>(server);
    if (!s || !out) return ZKMETAL_ERR_INVALID_INPUT;
    zk::ServerStats st;
    s->stats(&st);
    out->submitted = st.submitted;
    out->completed = st.completed;
    out->failed = st.failed;
    out->batches = st.batches;
    out->p50_ms = st.p50_ms;
    out->p99_ms = st.p99_ms;
    out->mean_ms = st.mean_ms;
    out->max_ms = st.max_ms;
    out->msms_per_s = st.msms_per_s;
    out->mean_batch = st.mean_batch;
    out->hbm_peak_bytes = st.hbm_peak;
    out->hbm_in_use_bytes = st.hbm_in_use;
    return ZKMETAL_SUCCESS;
}

void zkmsm_server_destroy(ZkmsmServer server) { delete static_cast<zk::Server*>(server); }

uint64_t zkmsm_server_footprint(uint32_t n) { return zk::Server::footprint(n); }

}  // extern "C"
