// qa_pool.cpp — the persistent host worker pool (see qa_pool.h).
#include "qa_pool.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace qapool {

namespace {
thread_local Pool *t_pool = nullptr;     // the pool of calls from this thread
thread_local bool t_in_task = false;     // this thread is running a pool task
}  // namespace

class Pool {
  public:
    explicit Pool(size_t threads) {
        for (size_t i = 1; i < threads; i++) th_.emplace_back([this] { work(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(size_t n, const std::function<void(size_t)> &f) {
        if (n == 0) return;
        if (t_in_task || th_.empty() || n == 1) {
            for (size_t i = 0; i < n; i++) f(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            n_ = n;
            next_.store(0);
            active_ = th_.size();
            gen_++;
        }
        cv_.notify_all();
        drain(f, n);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return active_ == 0; });
        job_ = nullptr;
    }

  private:
    void drain(const std::function<void(size_t)> &f, size_t n) {
        const bool was = t_in_task;
        t_in_task = true;
        for (size_t i; (i = next_.fetch_add(1)) < n;) f(i);
        t_in_task = was;
    }
    void work() {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::function<void(size_t)> *f = job_;
            const size_t n = n_;
            lk.unlock();
            drain(*f, n);
            lk.lock();
            if (--active_ == 0) done_.notify_all();
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)> *job_ = nullptr;
    size_t n_ = 0, active_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::vector<std::thread> th_;
};

Pool *create(size_t threads) { return new Pool(std::max<size_t>(threads, 1)); }
void destroy(Pool *p) { delete p; }

static Pool *default_pool() {
    // lives until the process ends (never destroyed: its threads wait idle)
    static Pool *p = new Pool(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    return p;
}

void run(size_t n, const std::function<void(size_t)> &f) {
    (t_pool ? t_pool : default_pool())->run(n, f);
}

Use::Use(Pool *p) : prev(t_pool) { t_pool = p; }
Use::~Use() { t_pool = prev; }

}  // namespace qapool
