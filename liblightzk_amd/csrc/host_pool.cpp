// Host worker pool (see host_pool.hpp). Workers: the OpenMP thread limit of
// the process (OMP_NUM_THREADS where set) minus the caller, at least one.
#include "host_pool.hpp"
#include <omp.h>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace hzk {
namespace {

struct Job {
    long n;
    const std::function<void(long)>* fn;
    std::atomic<long> next{0};
    long done = 0;   // under Pool::mu
    std::condition_variable finished;
};

thread_local bool t_in_pool = false;

struct Pool {
    std::mutex mu;
    std::condition_variable work;
    std::deque<std::shared_ptr<Job>> jobs;
    std::vector<std::thread> threads;
    bool stop = false;

    Pool() {
        const int n = std::max(1, omp_get_max_threads() - 1);
        for (int i = 0; i < n; ++i) threads.emplace_back([this] { worker(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        work.notify_all();
        for (auto& t : threads) t.join();
    }

    // indices of job j until none is left; returns how many this thread ran
    static long drain(Job& j) {
        long ran = 0;
        for (long i; (i = j.next.fetch_add(1)) < j.n; ++ran) (*j.fn)(i);
        return ran;
    }
    void finish(Job& j, long ran) {
        if (!ran) return;
        std::lock_guard<std::mutex> lk(mu);
        j.done += ran;
        if (j.done == j.n) j.finished.notify_all();
    }
    void worker() {
        t_in_pool = true;
        for (;;) {
            std::shared_ptr<Job> j;
            {
                std::unique_lock<std::mutex> lk(mu);
                work.wait(lk, [&] { return stop || !jobs.empty(); });
                if (stop && jobs.empty()) return;
                j = jobs.front();
                if (j->next.load() >= j->n) {   // every index handed out: off the queue
                    jobs.pop_front();
                    continue;
                }
            }
            finish(*j, drain(*j));
        }
    }
};

Pool& pool() {
    static Pool p;
    return p;
}

}  // namespace

void parallel_for(long n, const std::function<void(long)>& fn) {
    if (n <= 0) return;
    if (n == 1 || t_in_pool) {
        for (long i = 0; i < n; ++i) fn(i);
        return;
    }
    Pool& p = pool();
    auto j = std::make_shared<Job>();
    j->n = n;
    j->fn = &fn;
    {
        std::lock_guard<std::mutex> lk(p.mu);
        p.jobs.push_back(j);
    }
    p.work.notify_all();
    p.finish(*j, Pool::drain(*j));
    std::unique_lock<std::mutex> lk(p.mu);
    j->finished.wait(lk, [&] { return j->done == j->n; });
}

}  // namespace hzk
