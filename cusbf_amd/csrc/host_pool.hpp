// host_pool.hpp -- a persistent pool of host threads for memory copies that
// need more than one core's memory bandwidth (the request server's batch fills
// and answer copies, the pinned staging of the filter's host-memory forms).
// Internal to libcusbf_hip.so.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace cusbf_internal {

// The thread that submits a job works on it too and returns when every part
// is done; several threads may submit jobs concurrently.
class WorkerPool {
   public:
    explicit WorkerPool(unsigned threads) {
        for (unsigned i = 0; i < threads; ++i) workers_.emplace_back([this] { work(); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : workers_) t.join();
    }
    // Runs fn(i) for every i in [0, n).
    void run(size_t n, const std::function<void(size_t)>& fn) {
        if (n == 0) return;
        auto job = std::make_shared<Job>();
        job->fn = &fn;
        job->n = n;
        if (n > 1) {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(job);
        }
        cv_.notify_all();
        drain(*job);
        std::unique_lock<std::mutex> lk(job->mu);
        job->cv.wait(lk, [&] { return job->finished == job->n; });
    }

   private:
    struct Job {
        const std::function<void(size_t)>* fn = nullptr;
        size_t n = 0;
        std::atomic<size_t> next{0};
        size_t finished = 0;  // guarded by mu
        std::mutex mu;
        std::condition_variable cv;
    };
    static void drain(Job& job) {
        for (;;) {
            const size_t i = job.next.fetch_add(1);
            if (i >= job.n) return;
            (*job.fn)(i);
            std::lock_guard<std::mutex> lk(job.mu);
            if (++job.finished == job.n) job.cv.notify_all();
        }
    }
    void work() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
            if (jobs_.empty()) return;  // stopping
            std::shared_ptr<Job> job = jobs_.front();
            if (job->next.load() >= job->n) {  // every part taken: retire the job
                jobs_.pop_front();
                continue;
            }
            lk.unlock();
            drain(*job);
            lk.lock();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<Job>> jobs_;
    std::vector<std::thread> workers_;
    bool stop_ = false;
};


// The process-wide pool of the filter's host-memory forms (created on first use).
WorkerPool& host_pool();

// The host's first touch of a page of pinned memory costs a fault (~3 GB/s on
// one thread): touch (zero) `bytes` at `p` in 4 MiB parts on `pool`, so that
// later copies into the buffer run at memory speed.
inline void touch_pages(WorkerPool& pool, void* p, uint64_t bytes) {
    constexpr uint64_t kPart = 4ull << 20;
    const uint64_t parts = (bytes + kPart - 1) / kPart;
    pool.run(parts, [&](size_t i) {
        const uint64_t at = i * kPart;
        std::memset(static_cast<char*>(p) + at, 0, std::min(kPart, bytes - at));
    });
}

}  // namespace cusbf_internal
