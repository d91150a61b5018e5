// Request server: continuous batching of MSM commitment requests over a
// device-resident base table shared by several engine slots.
//
// Each slot is an Engine with its own HIP stream and workspace.  A slot's
// worker takes the queued requests in arrival order (up to max_batch) whose
// estimated device footprint fits the free part of the HBM budget, runs them
// as one batched launch sequence, and returns the budget when they finish.
// Requests arriving while slots are busy join the next batch; batches on
// different slots overlap on the device.  A request larger than the whole
// budget runs alone once nothing else is in flight.
#pragma once
#include <stdint.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace zk {

struct ServerConfig {
    uint32_t slots = 2;
    uint32_t max_batch = 64;
    uint64_t hbm_budget = 16ull << 30;
};

struct ServerStats {
    uint64_t submitted, completed, failed, batches;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double msms_per_s;  // completed / (last completion - first submission)
    double mean_batch;
    uint64_t hbm_peak, hbm_in_use;
};

class Server {
  public:
    // Serves the resident SRS of `owner` (whose mutex the caller holds):
    // the server attaches to it (the owner's srs_* calls that would replace
    // the table refuse until the server is destroyed) and holds a reference
    // (the owner lives until the server is gone).  Slot engines take the
    // owner's window tables.
    static int32_t create(Engine* owner, const ServerConfig& cfg, Server** out);
    ~Server();

    // copies the scalars (or, with copy == false, reads them in place: the
    // caller keeps them valid until the request is done); *ticket identifies it
    int32_t submit(const uint8_t* scalars, uint32_t n, uint64_t* ticket, bool copy = true);
    // block until done; writes the 96-byte result and the latency
    int32_t wait(uint64_t ticket, uint8_t* out96, double* latency_ms);
    int32_t poll(uint64_t ticket, uint8_t* out96, int32_t* done);
    void stats(ServerStats* s);
    static uint64_t footprint(uint32_t n);

  private:
    struct Request {
        uint64_t id;
        std::vector<uint8_t> owned;
        const uint8_t* scalars;
        uint32_t n;
        uint64_t bytes;
        std::chrono::steady_clock::time_point t_submit, t_done;
        uint8_t out[96];
        int32_t status = 0;
        bool done = false;
    };

    Server() = default;
    void worker(size_t slot);

    Engine* owner_ = nullptr;
    const void* table_ = nullptr;
    uint32_t table_n_ = 0;
    uint64_t table_shift_ = 0;  // the owner's window-shifted copies (stride, 0: none)
    ServerConfig cfg_;
    std::vector<Engine*> engines_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Request*> queue_;
    std::unordered_map<uint64_t, Request*> reqs_;
    uint64_t next_id_ = 1;
    uint64_t in_use_ = 0, peak_ = 0;
    bool stop_ = false;
    // statistics
    std::vector<double> lat_ms_;
    uint64_t submitted_ = 0, completed_ = 0, failed_ = 0, batches_ = 0, batched_reqs_ = 0;
    bool have_first_ = false;
    std::chrono::steady_clock::time_point first_submit_, last_done_;
};

}  // namespace zk
