// Host-side parallelism of the library: how many threads host work may use,
// and a memcpy split over a persistent set of worker threads (copies of
// trajectory chunks from pinned staging memory into the caller's array).
#pragma once

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace sodecl_amd {

// The CPUs this process may run on, at most OMP_NUM_THREADS when that is set
// (a machine's core count can exceed the share of it a process is given).
inline int host_threads()
{
    cpu_set_t set;
    int n = sched_getaffinity(0, sizeof set, &set) == 0 ? CPU_COUNT(&set) : 1;
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const int m = std::atoi(e);
        if (m > 0) n = std::min(n, m);
    }
    return std::max(n, 1);
}

// Worker threads that run the parts of one job at a time together with the
// calling thread.  Created on first use with host_threads() - 1 workers and
// never destroyed: between jobs they sleep on a condition variable, and the
// process's exit ends them (a static destructor could run while they wait).
class HostWorkers {
public:
    static HostWorkers& get()
    {
        static HostWorkers* w = new HostWorkers(host_threads() - 1);
        return *w;
    }
    int size() const { return (int)threads_.size() + 1; }

    // fn(k) for every k in [0, n), spread over the workers and the caller;
    // returns when all have finished.  Jobs from several callers run one
    // after another.
    void run(int n, const std::function<void(int)>& fn)
    {
        std::lock_guard<std::mutex> serial(job_mu_);
        auto job = std::make_shared<Job>();
        job->fn = &fn;
        job->n = n;
        job->left = n;
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = job;
            ++generation_;
        }
        cv_work_.notify_all();
        work(*job);
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return job->left == 0; });
    }

private:
    // One job: its own part counter, so a worker that wakes late for a job
    // already finished takes no part of the next one (fn is touched only
    // for parts below n, all of which end before run returns).
    struct Job {
        const std::function<void(int)>* fn = nullptr;
        int n = 0;
        int left = 0;   // parts not yet finished (guarded by mu_)
        std::atomic<int> next{0};
    };

    explicit HostWorkers(int workers)
    {
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
        for (auto& t : threads_) t.detach();
    }
    void loop()
    {
        unsigned long seen = 0;
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&] { return generation_ != seen; });
                seen = generation_;
                job = job_;
            }
            work(*job);
        }
    }
    // takes parts until none is left; the last one to finish wakes the caller
    void work(Job& job)
    {
        for (;;) {
            const int k = job.next.fetch_add(1);
            if (k >= job.n) return;
            (*job.fn)(k);
            std::lock_guard<std::mutex> lk(mu_);
            if (--job.left == 0) cv_done_.notify_all();
        }
    }

    std::mutex job_mu_, mu_;
    std::condition_variable cv_work_, cv_done_;
    std::vector<std::thread> threads_;
    std::shared_ptr<Job> job_;
    unsigned long generation_ = 0;
};

// Smallest slice of a parallel copy: a 16 MiB trajectory chunk gives every
// one of 16 host threads a slice (with 2 MiB slices only 8 took part, and
// the host copy set the pace of the whole transfer: C2's trajectory
// returned in 15.5-16.1 ms against 13.9-14.4 ms with 1 MiB).
constexpr size_t kCopySliceMin = (size_t)1 << 20;

// memcpy split into slices of at least kCopySliceMin over the host workers:
// a copy from pinned staging memory into a fresh caller array is bound by
// one core's copy rate and page faults, not by memory bandwidth.
inline void parallel_copy(void* dst, const void* src, size_t bytes)
{
    HostWorkers& w = HostWorkers::get();
    const int t = (int)std::min<size_t>((size_t)w.size(), std::max<size_t>(1, bytes / kCopySliceMin));
    if (t <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t slice = ((bytes + t - 1) / t + 63) & ~(size_t)63;   // t slices cover every byte
    w.run(t, [=](int k) {
        const size_t off = std::min(bytes, slice * (size_t)k), len = std::min(bytes - off, slice);
        if (len) std::memcpy((char*)dst + off, (const char*)src + off, len);
    });
}

}  // namespace sodecl_amd
