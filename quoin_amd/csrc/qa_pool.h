// qa_pool.h — the persistent host worker pool that runs the sequential coders
// (range coder, LZ, pco, the host half of the competition) one block or stream
// per task. The library keeps a default pool of min(16, hardware threads)
// threads; a QuoinCtx owns its own pool of the size it was created with
// (capi/src/lib.rs:161-224), used by the calls made through it.
#pragma once
#include <cstddef>
#include <functional>

namespace qapool {

class Pool;
Pool *create(size_t threads);  // threads >= 1 (the caller counts as one)
void destroy(Pool *p);

// f(i) for every i < n on the calling thread's pool (see Use), the caller
// working beside the pool's threads. A call made from inside a task runs its
// items inline. One job runs on a pool at a time; other callers wait.
void run(size_t n, const std::function<void(size_t)> &f);

// While it lives, calls from this thread use pool p (null: the default pool).
struct Use {
    explicit Use(Pool *p);
    ~Use();
    Use(const Use &) = delete;
    Use &operator=(const Use &) = delete;
    Pool *prev;
};

}  // namespace qapool
