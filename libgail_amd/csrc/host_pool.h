// A fixed set of host threads running queued tasks (the served form's Tier-2 pool, the
// encoder's frame copies, host rate control).
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gail_amd {

class Pool {
public:
    explicit Pool(int n) {
        for (int i = 0; i < n; ++i) threads_.emplace_back([this] { run(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;
    int size() const { return (int)threads_.size(); }
    // urgent tasks go to the front of the queue
    void post(std::function<void()> f, bool urgent = false) {
        {
            std::lock_guard<std::mutex> l(m_);
            if (urgent) q_.push_front(std::move(f));
            else q_.push_back(std::move(f));
        }
        cv_.notify_one();
    }

private:
    void run() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    std::vector<std::thread> threads_;
    bool stop_ = false;
};

// fn(chunk, slot) for every chunk in [0, chunks), spread over the calling thread and the
// pool's threads. slot (below pool.size() + 1) is the running thread's own for this call, so
// per-slot partial results need no locks. The caller takes chunks as well, so a call
// completes even while the pool's threads serve other work. fn must not throw.
template <class F>
void parallelChunks(Pool& pool, int chunks, const F& fn) {
    struct Job {
        std::atomic<int> next{0}, done{0}, slot{0};
        int chunks = 0;
        const F* fn = nullptr;
        std::mutex m;
        std::condition_variable cv;
    };
    auto job = std::make_shared<Job>();
    job->chunks = chunks;
    job->fn = &fn;
    // a pool thread that starts after the last chunk was taken finds none and never calls fn
    auto work = [](Job& j) {
        const int s = j.slot.fetch_add(1);
        int k = 0;
        for (int c; (c = j.next.fetch_add(1)) < j.chunks; ++k) (*j.fn)(c, s);
        if (k && j.done.fetch_add(k) + k == j.chunks) {
            std::lock_guard<std::mutex> l(j.m);
            j.cv.notify_all();
        }
    };
    for (int h = 0; h < std::min(pool.size(), chunks - 1); ++h) pool.post([job, work] { work(*job); });
    work(*job);
    std::unique_lock<std::mutex> l(job->m);
    job->cv.wait(l, [&] { return job->done.load() == job->chunks; });
}

}  // namespace gail_amd
