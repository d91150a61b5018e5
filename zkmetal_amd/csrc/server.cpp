// Request server (see server.hpp).
#include "server.hpp"

#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>

namespace zk {

// Device bytes a request may occupy while it runs: scalars, digit words,
// two sorted key/value copies and merge records (~300 B per point at the
// default window widths) plus its buckets.
uint64_t Server::footprint(uint32_t n) {
    const uint32_t c = default_window_bits(n);
    const uint64_t w = (128 + c - 1) / c;
    return (uint64_t)n * 32 + 2ull * n * w * 16 + w * (1ull << (c - 1)) * kBucketBytes + (1u << 20);
}

int32_t Server::create(Engine* owner, const ServerConfig& cfg, Server** out) {
    if (!out || !owner || cfg.slots == 0 || cfg.max_batch == 0) return kInvalidInput;
    *out = nullptr;
    const void* table = owner->srs_table();
    const uint32_t table_n = owner->srs_size();
    if (!table || table_n == 0) return kInvalidInput;
    Server* s = new Server();
    s->table_ = table;
    s->table_n_ = table_n;
    s->table_shift_ = owner->srs_shift_stride();
    s->cfg_ = cfg;
    for (uint32_t i = 0; i < cfg.slots; i++) {
        Engine* e = nullptr;
        int32_t rc = Engine::create(&e);
        if (rc != kOk) {
            delete s;
            return rc;
        }
        e->copy_tuning_from(*owner);
        s->engines_.push_back(e);
    }
    owner->attach_server();
    owner->retain();
    s->owner_ = owner;
    for (uint32_t i = 0; i < cfg.slots; i++) s->threads_.emplace_back(&Server::worker, s, i);
    *out = s;
    return kOk;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    for (auto& t : threads_) t.join();
    for (Engine* e : engines_) delete e;
    for (auto& kv : reqs_) delete kv.second;
    if (owner_) {
        {
            std::lock_guard<std::mutex> lk(owner_->mutex());
            owner_->detach_server();
        }
        Engine::release(owner_);
    }
}

int32_t Server::submit(const uint8_t* scalars, uint32_t n, uint64_t* ticket, bool copy) {
    if (!scalars || !ticket || n == 0 || n > table_n_) return kInvalidInput;
    Request* r = new Request();
    if (copy) {
        r->owned.assign(scalars, scalars + (uint64_t)n * 32);
        r->scalars = r->owned.data();
    } else {
        r->scalars = scalars;
    }
    r->n = n;
    r->bytes = footprint(n);
    r->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) {
            delete r;
            return kGpuError;
        }
        r->id = next_id_++;
        reqs_[r->id] = r;
        queue_.push_back(r);
        submitted_++;
        if (!have_first_) {
            have_first_ = true;
            first_submit_ = r->t_submit;
        }
        *ticket = r->id;
    }
    cv_work_.notify_one();
    return kOk;
}

void Server::worker(size_t slot) {
    Engine* eng = engines_[slot];
    std::vector<uint8_t> res;
    std::vector<uint32_t> counts;
    std::vector<const uint8_t*> parts;
    for (;;) {
        std::vector<Request*> batch;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                if (stop_) return;
                // admit queued requests in order while they fit the budget
                while (!queue_.empty() && batch.size() < cfg_.max_batch) {
                    Request* r = queue_.front();
                    const bool alone_ok = batch.empty() && in_use_ == 0;
                    if (in_use_ + bytes + r->bytes > cfg_.hbm_budget && !alone_ok) break;
                    queue_.pop_front();
                    batch.push_back(r);
                    bytes += r->bytes;
                }
                if (!batch.empty()) break;
                cv_work_.wait(lk);
            }
            in_use_ += bytes;
            peak_ = std::max(peak_, in_use_);
        }
        // requests left in the queue may fit another idle slot
        cv_work_.notify_one();
        counts.clear();
        parts.clear();
        for (Request* r : batch) {
            counts.push_back(r->n);
            parts.push_back(r->scalars);
        }
        res.assign(batch.size() * 96, 0);
        int32_t rc;
        {
            std::lock_guard<std::mutex> elk(eng->mutex());
            rc = eng->msm_table_parts(table_, table_n_, parts.data(), nullptr, counts.data(),
                                      (uint32_t)counts.size(), res.data(), 0, table_shift_);
        }
        const auto now = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t k = 0; k < batch.size(); k++) {
                Request* r = batch[k];
                memcpy(r->out, res.data() + k * 96, 96);
                r->status = rc;
                r->done = true;
                r->t_done = now;
                std::vector<uint8_t>().swap(r->owned);
                r->scalars = nullptr;
                if (rc == kOk) {
                    completed_++;
                    lat_ms_.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
                } else {
                    failed_++;
                }
            }
            batches_++;
            batched_reqs_ += batch.size();
            last_done_ = now;
            in_use_ -= bytes;
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

int32_t Server::wait(uint64_t ticket, uint8_t* out96, double* latency_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return kInvalidInput;
    Request* r = it->second;
    cv_done_.wait(lk, [&] { return r->done || stop_; });
    if (!r->done) return kGpuError;
    if (out96) memcpy(out96, r->out, 96);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    int32_t rc = r->status;
    reqs_.erase(it);
    delete r;
    return rc;
}

int32_t Server::poll(uint64_t ticket, uint8_t* out96, int32_t* done) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return kInvalidInput;
    Request* r = it->second;
    if (done) *done = r->done ? 1 : 0;
    if (!r->done) return kOk;
    if (out96) memcpy(out96, r->out, 96);
    int32_t rc = r->status;
    reqs_.erase(it);
    delete r;
    return rc;
}

void Server::stats(ServerStats* s) {
    std::lock_guard<std::mutex> lk(mu_);
    memset(s, 0, sizeof(*s));
    s->submitted = submitted_;
    s->completed = completed_;
    s->failed = failed_;
    s->batches = batches_;
    s->mean_batch = batches_ ? (double)batched_reqs_ / batches_ : 0;
    s->hbm_peak = peak_;
    s->hbm_in_use = in_use_;
    if (!lat_ms_.empty()) {
        std::vector<double> v = lat_ms_;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            size_t i = (size_t)std::min<double>(v.size() - 1, std::ceil(q * v.size()) - 1);
            return v[i];
        };
        s->p50_ms = pct(0.50);
        s->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s->mean_ms = sum / v.size();
        s->max_ms = v.back();
        double span = std::chrono::duration<double>(last_done_ - first_submit_).count();
        s->msms_per_s = span > 0 ? completed_ / span : 0;
    }
}

}  // namespace zk
