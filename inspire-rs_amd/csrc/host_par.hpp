// Host-side data parallelism for the host-buffer API: the reference's u64
// residues are narrowed into pinned staging (and responses widened back) on
// several threads, so the copies over PCIe, not one core, set the rate.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace inspire {

// Threads a host-buffer call uses: the machine's, at most 16.
inline unsigned host_threads() {
    static const unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return n;
}

// A fixed set of worker threads for par_ranges, started on first use and
// kept for the life of the process (the serving lanes call par_ranges for
// every batch; starting threads per call would cost more than the copies).
// Tasks are independent parts of callers' ranges; a caller waits only for
// its own parts, so concurrent callers share the workers.
class HostPool {
   public:
    static HostPool& get() {
        static HostPool* p = new HostPool(host_threads() - 1);  // never destroyed: workers outlive main
        return *p;
    }
    void submit(std::function<void()> task) {
        {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back(std::move(task));
        }
        cv_.notify_one();
    }
    size_t workers() const { return n_; }

   private:
    explicit HostPool(size_t n) : n_(n) {
        for (size_t i = 0; i < n; i++) std::thread([this] { loop(); }).detach();
    }
    void loop() {
        for (;;) {
            std::function<void()> t;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return !q_.empty(); });
                t = std::move(q_.front());
                q_.pop_front();
            }
            t();
        }
    }
    size_t n_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
};

// Runs f(begin, end) over contiguous ranges covering [0, n), each of at least
// `grain` items, on up to host_threads() threads (the caller runs one range,
// the pool's workers the others).  f must not throw.
template <class F>
void par_ranges(size_t n, size_t grain, F&& f) {
    if (!n) return;
    grain = std::max<size_t>(grain, 1);
    HostPool& pool = HostPool::get();
    const size_t parts = std::min<size_t>(pool.workers() + 1, (n + grain - 1) / grain);
    if (parts <= 1) {
        f((size_t)0, n);
        return;
    }
    const size_t step = (n + parts - 1) / parts;
    std::mutex m;
    std::condition_variable cv;
    size_t left = 0;  // parts on the workers, counted before any of them starts
    for (size_t p = 1; p < parts; p++)
        if (p * step < n) left++;
    for (size_t p = 1; p < parts; p++) {
        const size_t b = p * step, e = std::min(n, b + step);
        if (b >= e) continue;
        pool.submit([&, b, e] {
            f(b, e);
            std::lock_guard<std::mutex> lk(m);
            if (--left == 0) cv.notify_one();
        });
    }
    f((size_t)0, std::min(n, step));
    std::unique_lock<std::mutex> lk(m);
    cv.wait(lk, [&] { return left == 0; });
}

// dst[i] = (uint32_t)src[i]: residues (< 2^32) from the reference's u64 words.
inline void narrow_u64(uint32_t* dst, const uint64_t* src, size_t words) {
    for (size_t i = 0; i < words; i++) dst[i] = (uint32_t)src[i];
}

inline void widen_u32(uint64_t* dst, const uint32_t* src, size_t words) {
    for (size_t i = 0; i < words; i++) dst[i] = src[i];
}

}  // namespace inspire
