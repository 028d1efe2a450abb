// qa_server.cpp — the serving loop: continuous batching of compress /
// decompress jobs over the device-resident entry points, with admission by an
// HBM byte budget.
//
// Worker slots each own a QuoinGpu handle (its own HIP stream and scratch
// arenas, capped at budget / slots). A slot that is free takes the oldest
// waiting job of its class and every other waiting job of the same group
// (compress with the same config, or decompress) whose bytes fit the budget
// next to everything already admitted, and runs them as one batched launch
// sequence. With two or more slots, slot 0 serves only small requests (up to
// 1 MiB of values), so they never queue behind large ones. Batches of different slots
// run concurrently on their streams, so work of jobs that arrived while
// others were on the device overlaps with it. Columns are independent in a
// batch (every stage keys its state by column), so each job's bytes are the
// same alone and in any mix; a batch that reports an error is rerun one job
// at a time so each job gets its own return code.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/quoin.h"

namespace {
using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    QuoinJob job{};
    uint64_t bytes = 0;  // admission charge
    Clock::time_point t_submit, t_done;
    int rc = 0;
    uint64_t out_len = 0;
    bool done = false;
    bool small = false;  // uncompressed bytes <= kSmallBytes
};

// size class of a job: requests up to 1 MiB of values have a slot of their own
// (slot 0 when there are two or more), so they never wait behind large ones
constexpr uint64_t kSmallBytes = 1ull << 20;

size_t width_of(int32_t dtype) { return quoin_dtype_width(dtype); }

// bytes a job holds while admitted: its buffers plus the library scratch per
// value (encode state of the Fast pool is the largest, about 96 B per value)
uint64_t charge(const QuoinJob &j) {
    if (j.kind == QUOIN_JOB_COMPRESS)
        return j.column.n * width_of(j.column.dtype) + j.column.dst_cap + 96ull * j.column.n;
    return j.stream.src_len + j.stream.n * width_of(j.stream.dtype) + 48ull * j.stream.n;
}

bool same_group(const QuoinJob &a, const QuoinJob &b) {
    if (a.kind != b.kind) return false;
    if (a.kind == QUOIN_JOB_DECOMPRESS) return true;
    return a.cfg.level == b.cfg.level && a.cfg.block_size == b.cfg.block_size &&
           a.cfg.selection == b.cfg.selection && a.cfg.predictor_log2 == b.cfg.predictor_log2;
}
}  // namespace

struct QuoinServer {
    int device = 0;
    uint64_t budget = 0, inflight = 0;
    uint32_t max_batch = 256;
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;
    std::unordered_map<uint64_t, Req *> reqs;
    uint64_t next_ticket = 1, jobs = 0, batches = 0;
    bool stop = false;
    std::vector<QuoinGpu *> gpus;
    std::vector<std::thread> workers;
};

namespace {

void run_batch(QuoinGpu *g, std::vector<Req *> &batch) {
    const size_t n = batch.size();
    if (batch[0]->job.kind == QUOIN_JOB_COMPRESS) {
        std::vector<QuoinDeviceColumn> cols(n);
        std::vector<uint64_t> lens(n, 0);
        for (size_t i = 0; i < n; i++) cols[i] = batch[i]->job.column;
        int rc = quoin_gpu_compress(g, &batch[0]->job.cfg, cols.data(), n, lens.data(), nullptr);
        if (rc == QUOIN_OK || n == 1) {
            for (size_t i = 0; i < n; i++) {
                batch[i]->rc = rc;
                batch[i]->out_len = lens[i];
            }
            return;
        }
        for (size_t i = 0; i < n; i++) {
            uint64_t len = 0;
            batch[i]->rc = quoin_gpu_compress(g, &batch[i]->job.cfg, &cols[i], 1, &len, nullptr);
            batch[i]->out_len = len;
        }
        return;
    }
    std::vector<QuoinDeviceStream> st(n);
    for (size_t i = 0; i < n; i++) st[i] = batch[i]->job.stream;
    int rc = quoin_gpu_decompress(g, st.data(), n, nullptr);
    if (rc == QUOIN_OK || n == 1) {
        for (size_t i = 0; i < n; i++) batch[i]->rc = rc;
        return;
    }
    for (size_t i = 0; i < n; i++) batch[i]->rc = quoin_gpu_decompress(g, &st[i], 1, nullptr);
}

void worker(QuoinServer *S, size_t slot) {
    hipSetDevice(S->device);
    QuoinGpu *g = S->gpus[slot];
    const bool small_only = slot == 0 && S->gpus.size() >= 2;
    std::vector<Req *> batch;
    // this slot's next job: the oldest of its class (FIFO within the class),
    // provided it fits the budget beside everything admitted
    auto head = [&]() -> Req * {
        for (Req *r : S->pending) {
            if (small_only && !r->small) continue;
            return S->inflight + r->bytes <= S->budget ? r : nullptr;
        }
        return nullptr;
    };
    auto none_for_me = [&]() {
        for (Req *r : S->pending)
            if (!small_only || r->small) return false;
        return true;
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return (S->stop && none_for_me()) || head() != nullptr; });
            Req *h = head();
            if (!h) return;  // stopping with nothing of this slot's class queued
            batch.clear();
            const QuoinJob first = h->job;
            for (auto it = S->pending.begin();
                 it != S->pending.end() && batch.size() < S->max_batch;) {
                Req *r = *it;
                if ((!small_only || r->small) && same_group(r->job, first) &&
                    S->inflight + r->bytes <= S->budget) {
                    S->inflight += r->bytes;
                    batch.push_back(r);
                    it = S->pending.erase(it);
                } else {
                    ++it;
                }
            }
        }
        run_batch(g, batch);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            const auto now = Clock::now();
            for (Req *r : batch) {
                r->t_done = now;
                r->done = true;
                S->inflight -= r->bytes;
            }
            S->jobs += batch.size();
            S->batches++;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" QuoinServer *quoin_server_create(int device, uint64_t hbm_budget_bytes, uint32_t slots,
                                            uint32_t max_batch) {
    if (slots == 0) slots = 2;
    QuoinServer *S = new (std::nothrow) QuoinServer();
    if (!S) return nullptr;
    S->device = device < 0 ? 0 : device;
    S->budget = hbm_budget_bytes ? hbm_budget_bytes : (64ull << 30);
    S->max_batch = max_batch ? max_batch : 256;
    for (uint32_t i = 0; i < slots; i++) {
        // the budget is shared: each slot's arenas are capped at its share
        QuoinGpu *g = quoin_gpu_create(S->device, S->budget / slots);
        if (!g) {
            for (QuoinGpu *h : S->gpus) quoin_gpu_destroy(h);
            delete S;
            return nullptr;
        }
        S->gpus.push_back(g);
    }
    for (uint32_t i = 0; i < slots; i++) S->workers.emplace_back(worker, S, (size_t)i);
    return S;
}

extern "C" int quoin_server_submit(QuoinServer *S, const QuoinJob *job, uint64_t *ticket) {
    if (!S || !job || !ticket) return QUOIN_ERR_NULL;
    if (job->kind != QUOIN_JOB_COMPRESS && job->kind != QUOIN_JOB_DECOMPRESS)
        return QUOIN_ERR_UNSUPPORTED;
    Req *r = new (std::nothrow) Req();
    if (!r) return QUOIN_ERR_PANIC;
    r->job = *job;
    r->bytes = charge(*job);
    r->small = (job->kind == QUOIN_JOB_COMPRESS ? job->column.n * width_of(job->column.dtype)
                                                 : job->stream.n * width_of(job->stream.dtype)) <=
               kSmallBytes;
    // every slot's handle holds at most budget / slots of scratch
    if (r->bytes > S->budget / S->gpus.size()) {
        delete r;
        return QUOIN_ERR_BUFFER_TOO_SMALL;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete r;
            return QUOIN_ERR_PANIC;
        }
        r->ticket = S->next_ticket++;
        r->t_submit = Clock::now();
        S->reqs[r->ticket] = r;
        S->pending.push_back(r);
        *ticket = r->ticket;
    }
    S->cv_work.notify_all();  // the job's class decides which slots may take it
    return QUOIN_OK;
}

extern "C" int quoin_server_wait(QuoinServer *S, uint64_t ticket, uint64_t *out_len,
                                 double *latency_ms) {
    if (!S) return QUOIN_ERR_NULL;
    Req *r;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end()) return QUOIN_ERR_NULL;
        r = it->second;
        // the waiter claims the ticket before it sleeps: a second wait on it
        // (or after it) gets QUOIN_ERR_NULL, and no iterator is held across
        // the wait, during which submits may rehash the map
        S->reqs.erase(it);
        S->cv_done.wait(lk, [&] { return r->done; });
    }
    if (out_len) *out_len = r->out_len;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    int rc = r->rc;
    delete r;
    return rc;
}

extern "C" void quoin_server_stats(const QuoinServer *S, uint64_t *jobs, uint64_t *batches) {
    if (!S) return;
    std::lock_guard<std::mutex> lk(S->mu);
    if (jobs) *jobs = S->jobs;
    if (batches) *batches = S->batches;
}

extern "C" void quoin_server_destroy(QuoinServer *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (std::thread &t : S->workers) t.join();
    for (QuoinGpu *g : S->gpus) quoin_gpu_destroy(g);
    for (auto &kv : S->reqs) delete kv.second;
    delete S;
}
