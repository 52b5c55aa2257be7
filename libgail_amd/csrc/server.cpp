// Served form of libgail.so (gail_server_*, include/gail_amd.h): a request loop on one device
// for frames of many streams with mixed profiles and input formats.
//
// Device memory is held by jobs of the plugin: batches of frame slots of one class, 2K
// (2048 x 1080) or 4K (4096 x 2160), each slot room for one frame of its class in any input
// format. A scheduler thread admits waiting frames, the oldest first, into an idle job of
// their class (a 4K job also takes 2K frames), binds them, queues their inputs (DMA straight
// from pinned app memory when the frame already has the device layout, else a host copy on
// the pool) and launches the job; jobs run concurrently on their own streams. Jobs are made
// when frames wait and none of their class is idle, as far as the server's memory budget and
// the memory the device has free at that moment (less the headroom of pipeline.cpp:19,
// max(512 MiB, 6 %)) allow, up to max_jobs; idle jobs are retired to make room for the other
// class, and one at a time when the device runs short of memory (pipeline.cpp:105-138). A
// waiter thread per job collects its Tier-1 results and hands each frame to the host pool for
// rate control, Tier-2 and delivery; the job is idle again once all of its frames are done.
// A frame whose lambda falls below its early-termination threshold is queued again, first in
// line, to be coded without it (pipeline.cpp:194-218).
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "codestream.h"
#include "gail_amd.h"
#include "gail_hip_plugin.h"
#include "host_pool.h"
#include "stream.h"

using namespace gail_amd;

namespace {

using Clock = std::chrono::steady_clock;

int sfail(int status, const std::string& m) {
    lastErrorText() = m;
    return status;
}

double msSince(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

struct StreamState {
    std::unique_ptr<StreamSpec> spec;
    std::atomic<double> lastLambda{0.0};
};

struct Request {
    uint64_t id = 0;
    uint32_t stream = 0;
    StreamState* st = nullptr;
    gail_frame frame{};
    double slope = 0;
    bool autoSlope = false;
    bool recoded = false;
    bool onDevice = false;  // the planes are in device memory
    Clock::time_point tSubmit, tAdmit;
    int batchFrames = 0;
};

// Slot classes: a job's slots all take frames up to one of these sizes.
constexpr int kClasses = 2;
const GailHipSlot kClassSlot[kClasses] = {{2048, 1080, 1, 0}, {4096, 2160, 1, 0}};
const char* const kClassName[kClasses] = {"2K", "4K"};

// Launcher threads: each takes the next bound job, queues its uploads, waits for them and
// launches it. With one, jobs' uploads never share the link (two measured no faster).
constexpr int kLaunchers = 1;

struct JobState {
    int cls = 0;
    uint64_t bytes = 0;  // device memory the job holds
    GailHipBatch* batch = nullptr;
    std::vector<std::unique_ptr<Request>> reqs;
    std::vector<GailHipTier1> t1;
    int rcLeft = 0;
    bool busy = false;      // frames taken by the scheduler and not all delivered
    bool launched = false;  // waiter: a run is on the device
    bool retire = false;    // waiter: leave
    Clock::time_point lastUsed;
    std::thread waiter;
};

}  // namespace

struct gail_server {
    gail_server_config cfg{};
    const GailHipPlugin* P = nullptr;
    GailHipDevice* dev = nullptr;
    std::unique_ptr<Pool> pool;
    int slotsPerJob[kClasses] = {0, 0};
    uint64_t jobBytes[kClasses] = {0, 0}, slotBytes[kClasses] = {0, 0};

    std::mutex m;  // everything below
    std::condition_variable cv;
    std::deque<std::unique_ptr<Request>> pending;
    std::map<uint32_t, std::unique_ptr<StreamState>> streams;
    std::vector<std::unique_ptr<JobState>> jobs;
    uint64_t heldBytes = 0, peakBytes = 0;
    int jobsMade = 0, jobsRetired = 0, peakJobs = 0;
    Clock::time_point nextPressureCheck;
    bool stop = false;
    uint64_t submitted = 0, delivered = 0, failedCount = 0, recodedCount = 0, jobRuns = 0, jobFrames = 0;
    int inFlight = 0, maxInFlight = 0;
    bool haveFirst = false;
    Clock::time_point tFirst, tLast;
    std::vector<double> latency, queued;
    std::thread scheduler;
    // Jobs bound and filled, in the order the scheduler took them, with their frames' sources:
    // the launcher thread queues one job's uploads at a time, waits until they have arrived,
    // then launches that job, so each job's frames cross PCIe at the link's full rate and its
    // kernels start at once while the next job's frames follow, and no device queue ever
    // holds a wait on a copy.
    std::deque<std::pair<JobState*, std::vector<const void*>>> toLaunch;
    bool launcherStop = false;
    std::vector<std::thread> launchers;
    std::mutex deliverMutex;  // one callback at a time

    void deliver(Request& r, int status, const gail_packet* pkt, double lambda) {
        const Clock::time_point now = Clock::now();
        gail_server_result res{};
        res.id = r.id;
        res.stream = r.stream;
        res.status = status;
        res.packet = pkt;
        res.queued_ms = msSince(r.tSubmit, r.tAdmit);
        res.latency_ms = msSince(r.tSubmit, now);
        res.batch_frames = r.batchFrames;
        res.recoded = r.recoded ? 1 : 0;
        res.lambda = lambda;
        if (cfg.on_result) {
            std::lock_guard<std::mutex> l(deliverMutex);
            cfg.on_result(cfg.user, &res);
        }
        std::lock_guard<std::mutex> l(m);
        ++delivered;
        --inFlight;
        if (status != GAIL_OK) ++failedCount;
        if (r.recoded) ++recodedCount;
        latency.push_back(res.latency_ms);
        queued.push_back(res.queued_ms);
        tLast = now;
        cv.notify_all();
    }

    // Rate control, Tier-2 and delivery of frame i of job j (pool thread).
    void finishFrame(JobState& j, int i) {
        Request& r = *j.reqs[i];
        const StreamSpec& sp = *r.st->spec;
        bool requeue = false;
        try {
            const GailHipTier1& t = j.t1[i];
            const Tier1Ref ref{(const uint32_t*)t.info, t.passLen, t.passDist, t.data, (size_t)t.numInfo};
            RateResult rr;
            GailHipRate dr{};
            if (!cfg.host_rate_control && P->batch_rate(j.batch, i, &dr) == GAIL_HIP_OK) {
                // the device kept each block's chosen passes and only their bytes
                rr.passes.resize(ref.numBlocks);
                for (size_t b = 0; b < ref.numBlocks; ++b) rr.passes[b] = (uint8_t)ref.passes((int)b);
                for (int c = 0; c < 3; ++c) rr.lambdaComp[c] = dr.lambda[c];
                rr.distortion = dr.distortion;
            } else {
                rr = allocateRate(sp.L, ref, sp.limits);
            }
            const double lam = std::min({rr.lambdaComp[0], rr.lambdaComp[1], rr.lambdaComp[2]});
            if (r.autoSlope && r.slope > 0 && lam < r.slope) {
                requeue = true;  // passes the rate control keeps may have been skipped
            } else {
                std::vector<uint8_t> cs;
                buildCodestream(sp.L, ref, rr.passes.data(), &cs, nullptr, "hip");
                r.st->lastLambda.store(lam);
                gail_packet p{};
                p.data = cs.data();
                p.size = cs.size();
                p.index = (int64_t)r.id;
                p.opaque = r.frame.opaque;
                p.compliant = 1;
                if (cfg.check) p.compliant = checkDciCodestream(cs.data(), cs.size(), sp.eyeFps(), sp.fourk()).empty();
                const double mse = rr.distortion / ((double)sp.width * sp.height * 3);
                p.psnr_estimate = mse > 0 ? 10 * std::log10(4095.0 * 4095.0 / mse) : 99.0;
                deliver(r, GAIL_OK, &p, lam);
            }
        } catch (const std::exception& e) {
            lastErrorText() = e.what();
            deliver(r, GAIL_ERROR_FAILED, nullptr, 0);
        }
        std::lock_guard<std::mutex> l(m);
        if (requeue) {
            j.reqs[i]->slope = 0;
            j.reqs[i]->recoded = true;
            pending.push_front(std::move(j.reqs[i]));
        }
        if (--j.rcLeft == 0) {
            j.reqs.clear();
            j.busy = false;
        }
        cv.notify_all();
    }

    void waiterLoop(JobState* j) {
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m);
                cv.wait(l, [&] { return j->launched || j->retire; });
                if (!j->launched) return;
                j->launched = false;
            }
            const int n = (int)j->reqs.size();
            j->t1.assign(n, GailHipTier1{});
            if (P->batch_wait(j->batch, j->t1.data(), nullptr) != GAIL_HIP_OK) {
                const std::string why = P->last_error();
                for (auto& r : j->reqs) {
                    lastErrorText() = why;
                    deliver(*r, GAIL_ERROR_FAILED, nullptr, 0);
                }
                std::lock_guard<std::mutex> l(m);
                j->reqs.clear();
                j->busy = false;
                cv.notify_all();
                continue;
            }
            {
                std::lock_guard<std::mutex> l(m);
                j->rcLeft = n;
            }
            for (int i = 0; i < n; ++i) pool->post([this, j, i] { finishFrame(*j, i); });
        }
    }

    // Binds, fills and launches job j with the frames taken from the queue (scheduler).
    void admit(JobState& j) {
        const int n = (int)j.reqs.size();
        std::vector<GailHipLayout> layouts(n);
        std::vector<GailHipEncodeSettings> settings(n);
        for (int i = 0; i < n; ++i) {
            const StreamSpec& sp = *j.reqs[i]->st->spec;
            layouts[i] = sp.plainLayout;
            settings[i] = sp.settings;
            if (!cfg.host_rate_control) {  // rate control on the device, the stream's budgets
                settings[i].rcMaxTotal = sp.limits.maxTotal;
                settings[i].rcMaxComp = sp.limits.maxComp;
            }
        }
        auto failAll = [&](const std::string& why) { failJob(j, why); };
        if (P->batch_bind(j.batch, layouts.data(), settings.data(), n) != GAIL_HIP_OK) return failAll(P->last_error());
        // host copies of frames not already in the device layout, in parallel on the pool
        std::vector<const void*> src(n, nullptr);
        std::string stagingError;
        std::mutex fm;
        std::condition_variable fcv;
        int fills = 0;
        for (int i = 0; i < n; ++i) {
            Request& r = *j.reqs[i];
            const void* direct = r.st->spec->directInput(r.frame);
            if (direct) {
                src[i] = direct;
                continue;
            }
            if (r.onDevice) {  // checked at submit; the host cannot read device planes
                if (stagingError.empty()) stagingError = "a device-resident frame must have the device input layout";
                continue;
            }
            void* dst = P->batch_input(j.batch, i);
            if (!dst) {
                if (stagingError.empty()) stagingError = P->last_error();
                continue;
            }
            src[i] = dst;
            ++fills;
            // ahead of queued rate control: the device waits for these
            pool->post(
                [&, dst, rp = &r] {
                    rp->st->spec->fill(rp->frame, (uint8_t*)dst);
                    std::lock_guard<std::mutex> l(fm);
                    if (--fills == 0) fcv.notify_all();
                },
                true);
        }
        {
            std::unique_lock<std::mutex> l(fm);
            fcv.wait(l, [&] { return fills == 0; });
        }
        if (!stagingError.empty()) return failAll(stagingError);
        std::lock_guard<std::mutex> l(m);
        toLaunch.emplace_back(&j, std::move(src));
        cv.notify_all();
    }

    // Every frame of a job fails with the same error and the job is free again.
    void failJob(JobState& j, const std::string& why) {
        for (auto& r : j.reqs) {
            lastErrorText() = why;
            deliver(*r, GAIL_ERROR_FAILED, nullptr, 0);
        }
        std::lock_guard<std::mutex> l(m);
        j.reqs.clear();
        j.busy = false;
        cv.notify_all();
    }

    void launcherLoop() {
        for (;;) {
            std::pair<JobState*, std::vector<const void*>> next;
            {
                std::unique_lock<std::mutex> l(m);
                cv.wait(l, [&] { return launcherStop || !toLaunch.empty(); });
                if (toLaunch.empty()) return;
                next = std::move(toLaunch.front());
                toLaunch.pop_front();
            }
            JobState& j = *next.first;
            const int n = (int)j.reqs.size();
            bool ok = true;
            for (int i = 0; i < n && ok; ++i) {
                ok = P->batch_upload_from(j.batch, i, next.second[i]) == GAIL_HIP_OK;
                if (ok) P->batch_set_slope(j.batch, i, j.reqs[i]->slope);
            }
            if (ok && P->batch_inputs_ready) ok = P->batch_inputs_ready(j.batch) == GAIL_HIP_OK;
            if (ok) ok = P->batch_launch(j.batch, 0) == GAIL_HIP_OK;
            if (!ok) {
                failJob(j, P->last_error());
                continue;
            }
            std::lock_guard<std::mutex> l(m);
            j.launched = true;
            ++jobRuns;
            jobFrames += n;
            cv.notify_all();
        }
    }

    static int classOf(const Request& r) { return r.st->spec->fourk() ? 1 : 0; }

    // With m held: whether a new job of class k fits the server's budget and the memory the
    // device has free now, less the headroom for other users of the GPU.
    bool roomFor(int k) {
        if ((int)jobs.size() >= cfg.max_jobs) return false;
        if (cfg.memory_budget && heldBytes + jobBytes[k] > cfg.memory_budget) return false;
        GailHipMemory mem{};
        if (P->device_memory(dev, &mem) == GAIL_HIP_OK && mem.budgeted &&
            mem.available < jobBytes[k] + memoryHeadroom(mem.total))
            return false;
        return true;
    }

    // Whether a job of class k could ever be held: within the budget and, with nothing else
    // held, within the device's memory less the headroom.
    bool everFits(int k) {
        if (cfg.memory_budget && jobBytes[k] > cfg.memory_budget) return false;
        GailHipMemory mem{};
        if (P->device_memory(dev, &mem) == GAIL_HIP_OK && mem.budgeted && mem.total < jobBytes[k] + memoryHeadroom(mem.total))
            return false;
        return true;
    }

    // Makes a job of class k (m held on entry and exit, released while the plugin allocates).
    JobState* makeJob(std::unique_lock<std::mutex>& l, int k) {
        std::vector<GailHipSlot> slots(slotsPerJob[k], kClassSlot[k]);
        GailHipBatch* b = nullptr;
        l.unlock();
        const int st = P->slots_create(dev, slots.data(), slotsPerJob[k], nullptr, &b);
        l.lock();
        if (st != GAIL_HIP_OK) return nullptr;
        auto j = std::make_unique<JobState>();
        j->cls = k;
        j->bytes = jobBytes[k];
        j->batch = b;
        j->lastUsed = Clock::now();
        JobState* jp = j.get();
        j->waiter = std::thread([this, jp] { waiterLoop(jp); });
        jobs.push_back(std::move(j));
        heldBytes += jobBytes[k];
        peakBytes = std::max(peakBytes, heldBytes);
        peakJobs = std::max(peakJobs, (int)jobs.size());
        ++jobsMade;
        return jp;
    }

    // Frees idle job j (m held on entry and exit, released while its waiter leaves and the
    // plugin frees its memory).
    void retireJob(std::unique_lock<std::mutex>& l, JobState* j) {
        auto it = std::find_if(jobs.begin(), jobs.end(), [&](const std::unique_ptr<JobState>& p) { return p.get() == j; });
        std::unique_ptr<JobState> owned = std::move(*it);
        jobs.erase(it);
        heldBytes -= owned->bytes;
        ++jobsRetired;
        owned->retire = true;
        cv.notify_all();
        l.unlock();
        owned->waiter.join();
        P->batch_destroy(owned->batch);
        owned.reset();
        l.lock();
    }

    JobState* idleJob(int k) {
        for (auto& j : jobs)
            if (!j->busy && j->cls == k) return j.get();
        return nullptr;
    }

    // The least recently used idle job not of class `keep` (-1: any class).
    JobState* idleToRetire(int keep) {
        JobState* best = nullptr;
        for (auto& j : jobs)
            if (!j->busy && j->cls != keep && (!best || j->lastUsed < best->lastUsed)) best = j.get();
        return best;
    }

    // Fails the waiting frames of class k (m held on entry and exit).
    void failClass(std::unique_lock<std::mutex>& l, int k, const std::string& why) {
        std::vector<std::unique_ptr<Request>> gone;
        for (auto it = pending.begin(); it != pending.end();) {
            if (classOf(**it) == k) {
                gone.push_back(std::move(*it));
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        l.unlock();
        for (auto& r : gone) {
            r->tAdmit = Clock::now();
            lastErrorText() = why;
            deliver(*r, GAIL_ERROR_NOMEM, nullptr, 0);
        }
        l.lock();
    }

    // A job for the waiting frames, reserved (busy), or nullptr when every usable job is busy
    // and no other can be made now. Classes are served in the order of their oldest frame.
    JobState* chooseJob(std::unique_lock<std::mutex>& l) {
        for (;;) {
            int order[kClasses], n = 0;
            for (const auto& r : pending) {
                const int k = classOf(*r);
                if (std::find(order, order + n, k) == order + n) order[n++] = k;
                if (n == kClasses) break;
            }
            bool again = false;
            for (int o = 0; o < n && !again; ++o) {
                const int k = order[o];
                if (!everFits(k) && !(k == 0 && everFits(1))) {
                    failClass(l, k, std::string("a job of ") + kClassName[k] + " frame slots exceeds the memory budget");
                    again = true;
                    continue;
                }
                JobState* j = idleJob(k);
                if (!j && roomFor(k)) j = makeJob(l, k);
                if (!j && k == 0) j = idleJob(1);  // a 4K job's slots take 2K frames
                if (j) {
                    j->busy = true;
                    return j;
                }
                // make room: the least recently used idle job of another class goes
                if (JobState* r = idleToRetire(k)) {
                    retireJob(l, r);
                    again = true;
                } else if (jobs.empty()) {
                    failClass(l, k, std::string("no device memory for a job of ") + kClassName[k] + " frame slots");
                    again = true;
                }
            }
            if (!again) return nullptr;
            if (pending.empty()) return nullptr;
        }
    }

    // Under memory pressure from other users of the device (free memory below half the
    // headroom) an idle job is given up, one at a time, the last one kept (pipeline.cpp:121-138).
    void relievePressure(std::unique_lock<std::mutex>& l) {
        const Clock::time_point now = Clock::now();
        if (now < nextPressureCheck || jobs.size() <= 1) return;
        nextPressureCheck = now + std::chrono::milliseconds(500);
        GailHipMemory mem{};
        if (P->device_memory(dev, &mem) != GAIL_HIP_OK || !mem.budgeted || mem.available >= memoryHeadroom(mem.total) / 2)
            return;
        if (JobState* j = idleToRetire(-1)) {
            retireJob(l, j);
            nextPressureCheck = Clock::now() + std::chrono::milliseconds(1000);  // let the freed memory show
        }
    }

    void schedulerLoop() {
        std::unique_lock<std::mutex> l(m);
        for (;;) {
            if (stop && pending.empty()) return;
            if (pending.empty()) {
                cv.wait(l);
                continue;
            }
            relievePressure(l);
            JobState* j = chooseJob(l);
            if (!j) {
                // every usable job busy and none can be made: a job going idle, a new frame or
                // memory coming free (checked again at least every 100 ms) wakes this
                cv.wait_for(l, std::chrono::milliseconds(100));
                continue;
            }
            const int cap = slotsPerJob[j->cls];
            auto fits = [&](const Request& r) { return classOf(r) <= j->cls; };
            auto waiting = [&] {
                int c = 0;
                for (const auto& r : pending) c += fits(*r) ? 1 : 0;
                return c;
            };
            // batching window: a part-full job waits for more frames, up to the window counted
            // from the arrival of the oldest frame it can take
            if (cfg.batch_window_us > 0 && !stop && waiting() < cap) {
                Clock::time_point oldest = Clock::now();
                for (const auto& r : pending)
                    if (fits(*r)) {
                        oldest = r->tSubmit;
                        break;
                    }
                cv.wait_until(l, oldest + std::chrono::microseconds(cfg.batch_window_us),
                              [&] { return stop || waiting() >= cap; });
            }
            const Clock::time_point now = Clock::now();
            for (auto it = pending.begin(); it != pending.end() && (int)j->reqs.size() < cap;) {
                if (!fits(**it)) {
                    ++it;
                    continue;
                }
                std::unique_ptr<Request> r = std::move(*it);
                it = pending.erase(it);
                r->tAdmit = now;
                if (r->autoSlope && !r->recoded) r->slope = r->st->spec->slopeFactor * r->st->lastLambda.load();
                j->reqs.push_back(std::move(r));
            }
            if (j->reqs.empty()) {  // the frames it was reserved for went to another job meanwhile
                j->busy = false;
                continue;
            }
            for (auto& r : j->reqs) r->batchFrames = (int)j->reqs.size();
            j->lastUsed = now;
            l.unlock();
            admit(*j);
            l.lock();
        }
    }

    void shutdown() {
        {
            std::lock_guard<std::mutex> l(m);
            stop = true;
        }
        cv.notify_all();
        if (scheduler.joinable()) scheduler.join();
        std::unique_lock<std::mutex> l(m);
        cv.wait(l, [&] {
            for (auto& j : jobs)
                if (j->busy) return false;
            return true;
        });
        launcherStop = true;
        for (auto& j : jobs) j->retire = true;
        cv.notify_all();
        l.unlock();
        for (auto& t : launchers)
            if (t.joinable()) t.join();
        for (auto& j : jobs)
            if (j->waiter.joinable()) j->waiter.join();
        pool.reset();
        for (auto& j : jobs)
            if (j->batch) P->batch_destroy(j->batch);
        jobs.clear();
        if (dev) P->device_destroy(dev);
        dev = nullptr;
    }
};

extern "C" {

GAIL_API void gail_server_config_init(gail_server_config* c, size_t size) {
    if (!c || size < sizeof(size_t)) return;
    gail_server_config d{};
    d.struct_size = size;
    d.slots_per_job = 8;
    d.max_jobs = 4;
    std::memset(c, 0, size);
    std::memcpy(c, &d, std::min(size, sizeof d));
}

GAIL_API int gail_server_create(const gail_server_config* c, gail_server** out) {
    if (!c || !out) return sfail(GAIL_ERROR_INVALID, "null argument");
    *out = nullptr;
    auto s = std::make_unique<gail_server>();
    gail_server_config_init(&s->cfg, sizeof s->cfg);
    std::memcpy(&s->cfg, c, std::min(c->struct_size ? c->struct_size : sizeof *c, sizeof s->cfg));
    gail_server_config& cfg = s->cfg;
    if (cfg.slots_per_job <= 0) cfg.slots_per_job = 8;
    if (cfg.max_jobs <= 0) cfg.max_jobs = 4;
    if (cfg.slots_per_job > 256 || cfg.max_jobs > 64) return sfail(GAIL_ERROR_INVALID, "slots_per_job or max_jobs too large");
    const PluginHandle& ph = loadPlugin();
    if (!ph.table) return sfail(GAIL_ERROR_UNSUPPORTED, ph.error);
    s->P = ph.table;
    const GailHipPlugin* P = s->P;
    const int n = P->device_count();
    if (n <= 0) return sfail(GAIL_ERROR_UNSUPPORTED, "no HIP device");
    if (cfg.device < 0 || cfg.device >= n) return sfail(GAIL_ERROR_INVALID, "no such device");
    if (P->device_usable(cfg.device) != GAIL_HIP_OK) return sfail(GAIL_ERROR_UNSUPPORTED, P->last_error());
    if (P->device_create(cfg.device, nullptr, nullptr, &s->dev) != GAIL_HIP_OK) return sfail(GAIL_ERROR_FAILED, P->last_error());
    // jobs of each slot class: as many slots as asked, fewer when one job would not fit the
    // memory budget (the server's own, or what the device has free now less the headroom)
    uint64_t budget = cfg.memory_budget;
    GailHipMemory mem{};
    if (P->device_memory(s->dev, &mem) == GAIL_HIP_OK && mem.budgeted) {
        const uint64_t head = memoryHeadroom(mem.total), usable = mem.available > head ? mem.available - head : 0;
        budget = budget ? std::min(budget, usable) : usable;
    }
    for (int k = 0; k < kClasses; ++k) {
        s->slotBytes[k] = P->slots_bytes(s->dev, &kClassSlot[k], 1, nullptr);
        int n = cfg.slots_per_job;
        std::vector<GailHipSlot> slots(n, kClassSlot[k]);
        uint64_t bytes = P->slots_bytes(s->dev, slots.data(), n, nullptr);
        while (n > 1 && budget && bytes > budget) bytes = P->slots_bytes(s->dev, slots.data(), --n, nullptr);
        s->slotsPerJob[k] = n;
        s->jobBytes[k] = bytes;
    }
    if (budget && s->jobBytes[0] > budget) {
        s->shutdown();
        return sfail(GAIL_ERROR_NOMEM, "the device memory budget holds no frame slot");
    }
    int threads = cfg.cpu_threads;
    if (threads <= 0) {  // the cores this process may run on (usableCpuCount, cpu_backend.cpp:180-192)
        cpu_set_t set;
        CPU_ZERO(&set);
        const int cores = sched_getaffinity(0, sizeof set, &set) == 0 ? CPU_COUNT(&set)
                                                                      : (int)std::thread::hardware_concurrency();
        threads = std::max(2, cores - 1);
    }
    s->pool = std::make_unique<Pool>(threads);
    for (int i = 0; i < kLaunchers; ++i) s->launchers.emplace_back([sp = s.get()] { sp->launcherLoop(); });
    s->scheduler = std::thread([sp = s.get()] { sp->schedulerLoop(); });
    *out = s.release();
    return GAIL_OK;
}

GAIL_API int gail_server_submit(gail_server* s, uint64_t id, uint32_t stream, const gail_config* c, const gail_frame* f,
                                double slope) {
    if (!s || !c || !f) return sfail(GAIL_ERROR_INVALID, "null argument");
    auto r = std::make_unique<Request>();
    r->id = id;
    r->stream = stream;
    r->frame = *f;
    r->tSubmit = Clock::now();
    {
        std::unique_lock<std::mutex> l(s->m);
        if (s->stop) return sfail(GAIL_ERROR_STATE, "the server is shutting down");
        auto& slot = s->streams[stream];
        if (!slot) {
            l.unlock();
            auto fresh = std::make_unique<StreamState>();
            try {
                fresh->spec = StreamSpec::make(*c);
            } catch (const ApiError& e) {
                return sfail(e.status, e.what());
            } catch (const std::exception& e) {
                return sfail(GAIL_ERROR_INVALID, e.what());
            }
            l.lock();
            auto& again = s->streams[stream];
            if (!again) again = std::move(fresh);
        }
        StreamState* sp = s->streams[stream].get();
        const gail_config& k = sp->spec->cfg;
        if (k.profile != c->profile || k.container != c->container || k.input.format != c->input.format ||
            k.fps != c->fps || k.max_mbps != c->max_mbps)
            return sfail(GAIL_ERROR_INVALID, "a stream's frames must keep its configuration");
        std::string why;
        const int chk = sp->spec->checkFrame(*f, why);
        if (chk != GAIL_OK) return sfail(chk, why);
        r->onDevice = s->P->is_device_pointer(f->data[0]) == 1;
        if (r->onDevice && !sp->spec->directInput(*f))
            return sfail(GAIL_ERROR_INVALID, "a device-resident frame must have the device input layout "
                                             "(planes back to back, no padding, the container size for X'Y'Z')");
        r->st = sp;
        r->autoSlope = slope < 0;
        r->slope = slope < 0 ? 0 : slope;
        if (!s->haveFirst) {
            s->haveFirst = true;
            s->tFirst = r->tSubmit;
        }
        ++s->submitted;
        s->maxInFlight = std::max(s->maxInFlight, ++s->inFlight);
        s->pending.push_back(std::move(r));
    }
    s->cv.notify_all();
    return GAIL_OK;
}

GAIL_API int gail_server_drain(gail_server* s) {
    if (!s) return sfail(GAIL_ERROR_INVALID, "null argument");
    std::unique_lock<std::mutex> l(s->m);
    s->cv.wait(l, [&] { return s->delivered >= s->submitted; });
    return GAIL_OK;
}

GAIL_API int gail_server_get_stats(gail_server* s, gail_server_stats* out) {
    if (!s || !out) return sfail(GAIL_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> l(s->m);
    gail_server_stats r{};
    r.frames = s->delivered;
    r.failed = s->failedCount;
    r.recoded = s->recodedCount;
    r.seconds = s->haveFirst ? msSince(s->tFirst, s->tLast) / 1000.0 : 0;
    r.fps = r.seconds > 0 ? (double)(s->delivered - s->failedCount) / r.seconds : 0;
    auto pct = [](std::vector<double> v, double q) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        const double pos = q * (double)(v.size() - 1);
        const size_t i = (size_t)pos;
        const double f = pos - (double)i;
        return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    r.p50_ms = pct(s->latency, 0.5);
    r.p99_ms = pct(s->latency, 0.99);
    r.max_ms = s->latency.empty() ? 0 : *std::max_element(s->latency.begin(), s->latency.end());
    r.queued_p50_ms = pct(s->queued, 0.5);
    r.queued_p99_ms = pct(s->queued, 0.99);
    r.mean_batch = s->jobRuns ? (double)s->jobFrames / (double)s->jobRuns : 0;
    r.jobs = (int)s->jobs.size();
    r.slots_per_job = s->slotsPerJob[1];
    r.max_frames_in_flight = s->maxInFlight;
    r.job_bytes = s->heldBytes;
    r.slot_bytes = s->slotBytes[1];
    for (auto& j : s->jobs) (j->cls ? r.jobs_4k : r.jobs_2k)++;
    r.jobs_made = s->jobsMade;
    r.jobs_retired = s->jobsRetired;
    r.max_jobs = s->peakJobs;
    r.slots_per_job_2k = s->slotsPerJob[0];
    r.peak_job_bytes = s->peakBytes;
    r.slot_bytes_2k = s->slotBytes[0];
    *out = r;
    return GAIL_OK;
}

GAIL_API void* gail_host_alloc(size_t bytes) {
    const PluginHandle& ph = loadPlugin();
    if (!ph.table) {
        lastErrorText() = ph.error;
        return nullptr;
    }
    void* p = ph.table->host_alloc(bytes);
    if (!p) lastErrorText() = "out of pinned host memory";
    return p;
}

GAIL_API void* gail_device_alloc(size_t bytes) {
    const PluginHandle& ph = loadPlugin();
    if (!ph.table) {
        lastErrorText() = ph.error;
        return nullptr;
    }
    GailHipDevice* d = nullptr;
    if (ph.table->device_count() <= 0 || ph.table->device_create(0, nullptr, nullptr, &d) != GAIL_HIP_OK) {
        lastErrorText() = "no HIP device";
        return nullptr;
    }
    void* p = ph.table->device_alloc(d, bytes);
    if (!p) lastErrorText() = ph.table->last_error();
    ph.table->device_destroy(d);
    return p;
}

GAIL_API void gail_device_free(void* p) {
    const PluginHandle& ph = loadPlugin();
    if (p && ph.table) ph.table->device_free(p);
}

GAIL_API int gail_copy(void* dst, const void* src, size_t bytes) {
    const PluginHandle& ph = loadPlugin();
    if (!ph.table) return sfail(GAIL_ERROR_UNSUPPORTED, ph.error);
    if (ph.table->copy(dst, src, bytes) != GAIL_HIP_OK) return sfail(GAIL_ERROR_FAILED, ph.table->last_error());
    return GAIL_OK;
}

GAIL_API void gail_host_free(void* p) {
    const PluginHandle& ph = loadPlugin();
    if (p && ph.table) ph.table->host_free(p);
}

GAIL_API void gail_server_destroy(gail_server* s) {
    if (!s) return;
    gail_server_drain(s);
    s->shutdown();
    delete s;
}

}  // extern "C"
