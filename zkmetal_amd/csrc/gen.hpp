// Seeded synthetic scalars (host).  Element i of a stream draws 64-bit words
// j = 0..63 from the splitmix64 output at counter i*64 + j + 1, so every
// element can be generated independently and in parallel.
//   dist 0: uniform in Fr (254-bit candidates, rejection sampling below r)
//   dist 1: witness-like mix: 40% zero, 30% uniform below 2^16,
//           10% one of {1, r-1}, 20% uniform in Fr
//   dist 2: uniform 256-bit integers (may exceed r)
#pragma once
#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace zkh {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

inline uint64_t gen_word(uint64_t seed, uint64_t i, uint64_t j) {
    return mix64(seed + 0x9e3779b97f4a7c15ULL * (i * 64 + j + 1));
}

static const uint64_t kR[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL,
                               0xb85045b68181585dULL, 0x30644e72e131a029ULL};

inline bool lt_r(const uint64_t v[4]) {
    for (int k = 3; k >= 0; k--)
        if (v[k] != kR[k]) return v[k] < kR[k];
    return false;
}

inline void uniform_fr(uint64_t seed, uint64_t i, uint64_t out[4]) {
    for (int a = 0; a < 16; a++) {
        for (int t = 0; t < 4; t++) out[t] = gen_word(seed, i, 4 * a + t);
        out[3] &= (1ULL << 62) - 1;
        if (lt_r(out)) return;
    }
    // 16 rejections in a row (probability ~1e-10): reduce the last candidate
    while (!lt_r(out)) {
        unsigned __int128 br = 0;
        for (int t = 0; t < 4; t++) {
            unsigned __int128 x = (unsigned __int128)out[t] - kR[t] - br;
            out[t] = (uint64_t)x;
            br = (x >> 127) & 1;
        }
    }
}

inline void gen_scalar(uint64_t seed, uint64_t i, int dist, uint64_t out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    if (dist == 0) {
        uniform_fr(seed, i, out);
    } else if (dist == 2) {
        for (int t = 0; t < 4; t++) out[t] = gen_word(seed, i, t);
    } else {
        uint64_t cls = gen_word(seed, i, 60) % 100;
        if (cls < 40) {
        } else if (cls < 70) {
            out[0] = gen_word(seed, i, 61) & 0xffff;
        } else if (cls < 80) {
            if (gen_word(seed, i, 62) & 1) {
                for (int t = 0; t < 4; t++) out[t] = kR[t];
                out[0] -= 1;
            } else {
                out[0] = 1;
            }
        } else {
            uniform_fr(seed, i, out);
        }
    }
}

// n scalars as 32-byte little-endian records, generated with up to `threads`
// host threads (0 = hardware concurrency, capped at 16)
inline void gen_scalars(uint64_t seed, uint64_t n, int dist, uint8_t* out, int threads = 0) {
    int nt = threads > 0 ? threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    auto work = [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) gen_scalar(seed, i, dist, reinterpret_cast<uint64_t*>(out + 32 * i));
    };
    std::vector<std::thread> th;
    uint64_t chunk = (n + nt - 1) / nt;
    for (int t = 1; t < nt; t++) {
        uint64_t b = std::min(n, t * chunk), e = std::min(n, b + chunk);
        th.emplace_back(work, b, e);
    }
    work(0, std::min(n, chunk));
    for (auto& x : th) x.join();
}

}  // namespace zkh
