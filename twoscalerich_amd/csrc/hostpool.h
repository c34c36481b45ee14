// hostpool.h - host threads for the per-scenario work of a scenario batch.
//
// A step of a batch of S scenarios evaluates, per scenario, small dense functions of its
// Krylov Hessenberg matrix on the host (phi_1 for the stopping rule and the step, phi_3
// for the error estimate; O(m^3) each, m <= 128).  Scenario q's evaluation reads and
// writes only scenario q's state, so the batch's results do not depend on which thread
// evaluates which scenario or in what order; spreading them over host threads shortens
// the device's wait for the host.  One pool per process, created on first use; its size
// is TSR_HOST_THREADS if set (1 = every scenario on the calling thread), else the CPUs
// this process may run on, at most 16.  set_limit(1) later runs every batch on the
// calling thread, set_limit(0) uses the pool again.  The worker threads make no HIP calls.
//
// The calling thread starts on the batch at once and never waits for a worker to wake:
// indices are claimed from a counter tagged with the batch's generation (a worker that
// wakes after its batch is over claims nothing and never touches the batch's function),
// and the call returns when every index has completed.  Workers spin briefly before
// sleeping, since a step's batches come in quick succession.
#pragma once

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace tsr {

class HostPool {
  public:
    static HostPool& get()
    {
        // never destroyed: the workers are parked on a condition variable at exit
        static HostPool* p = new HostPool(default_threads());
        return *p;
    }
    int threads() const { return limit_.load() == 1 ? 1 : nthr_; }
    void set_limit(int n) { limit_.store(n == 1 ? 1 : 0); }

    // f(i) for every i in [0, n), the calling thread taking part; returns when all are done.
    // Calls from several threads are served one batch at a time.
    void run(int n, const std::function<void(int)>& f)
    {
        if (threads() == 1 || n < 2) {
            for (int i = 0; i < n; ++i) f(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_m_);
        const long long g = gen_.load() + 1;
        {
            std::lock_guard<std::mutex> l(m_);
            job_ = &f;
            n_ = n;
            done_.store(0);
            claim_.store((unsigned long long)g << 32);   // generation g, next index 0
            gen_.store(g, std::memory_order_release);
        }
        cv_.notify_all();
        work(g, &f, n);
        while (done_.load(std::memory_order_acquire) < n) std::this_thread::yield();
    }

  private:
    explicit HostPool(int nthr) : nthr_(std::max(1, nthr))
    {
        for (int t = 1; t < nthr_; ++t) std::thread([this] { loop(); }).detach();
    }
    static int default_threads()
    {
        if (const char* e = std::getenv("TSR_HOST_THREADS")) {
            const int v = std::atoi(e);
            if (v >= 1) return std::min(v, 64);
        }
        int n = (int)std::thread::hardware_concurrency();
        cpu_set_t cs;
        CPU_ZERO(&cs);
        if (sched_getaffinity(0, sizeof(cs), &cs) == 0) n = CPU_COUNT(&cs);
        return std::max(1, std::min(n, 16));
    }
    // the next index of batch g, or -1 when it has none left (or a later batch began)
    int claim(long long g, int n)
    {
        unsigned long long v = claim_.load();
        for (;;) {
            if ((unsigned)(v >> 32) != (unsigned)g) return -1;   // generation tag: g's low 32 bits
            const int i = (int)(v & 0xffffffffull);
            if (i >= n) return -1;
            if (claim_.compare_exchange_weak(v, v + 1)) return i;
        }
    }
    void work(long long g, const std::function<void(int)>* f, int n)
    {
        for (int i = claim(g, n); i >= 0; i = claim(g, n)) {
            (*f)(i);
            done_.fetch_add(1, std::memory_order_release);
        }
    }
    void loop()
    {
        long long seen = 0;
        for (;;) {
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen &&
                   std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200))
                std::this_thread::yield();
            long long g;
            const std::function<void(int)>* f;
            int n;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return gen_.load() != seen; });
                g = gen_.load();
                f = job_;
                n = n_;
            }
            seen = g;
            work(g, f, n);
        }
    }

    const int nthr_;
    std::mutex run_m_, m_;
    std::condition_variable cv_;
    const std::function<void(int)>* job_ = nullptr;
    int n_ = 0;
    std::atomic<long long> gen_{0};
    std::atomic<unsigned long long> claim_{0};
    std::atomic<int> done_{0};
    std::atomic<int> limit_{0};
};

}  // namespace tsr
