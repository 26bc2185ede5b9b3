// Host-side data parallelism for the host-buffer API: the reference's u64
// residues are narrowed into pinned staging (and responses widened back) on
// several threads, so the copies over PCIe, not one core, set the rate.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

namespace inspire {

// Threads a host-buffer call uses: the machine's, at most 16.
inline unsigned host_threads() {
    static const unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return n;
}

// Runs f(begin, end) over contiguous ranges covering [0, n), each of at least
// `grain` items, on up to host_threads() threads (the caller runs one range).
// f must not throw.
template <class F>
void par_ranges(size_t n, size_t grain, F&& f) {
    if (!n) return;
    grain = std::max<size_t>(grain, 1);
    const size_t parts = std::min<size_t>(host_threads(), (n + grain - 1) / grain);
    if (parts <= 1) {
        f((size_t)0, n);
        return;
    }
    const size_t step = (n + parts - 1) / parts;
    std::vector<std::thread> th;
    th.reserve(parts - 1);
    for (size_t p = 1; p < parts; p++) {
        const size_t b = p * step, e = std::min(n, b + step);
        if (b < e) th.emplace_back([&f, b, e] { f(b, e); });
    }
    f((size_t)0, std::min(n, step));
    for (auto& t : th) t.join();
}

// dst[i] = (uint32_t)src[i]: residues (< 2^32) from the reference's u64 words.
inline void narrow_u64(uint32_t* dst, const uint64_t* src, size_t words) {
    for (size_t i = 0; i < words; i++) dst[i] = (uint32_t)src[i];
}

inline void widen_u32(uint64_t* dst, const uint32_t* src, size_t words) {
    for (size_t i = 0; i < words; i++) dst[i] = src[i];
}

}  // namespace inspire
