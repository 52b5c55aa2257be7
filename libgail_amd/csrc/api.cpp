// The encoder half of the GAIL library API (include/gail_amd.h) on the HIP backend plugin,
// with a single-device frame loop in place of the reference's multi-device pull scheduler
// (pipeline.cpp:42-292): frames handed over by the app go into a bounded queue; one worker
// thread per frame slot (an encoder of the plugin) codes Tier-1 on the GPU, then runs rate
// control and Tier-2 on the CPU; a writer thread delivers codestreams in send order to the
// waiter, the callback or the receive queue (gail_api.cpp:70-134, 323-417, 566-806).
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "codestream.h"
#include "gail_amd.h"
#include "gail_hip_plugin.h"
#include "host_pool.h"
#include "layout.h"
#include "stream.h"

using namespace gail_amd;

namespace {

int fail(int status, const std::string& msg) {
    lastErrorText() = msg;
    return status;
}

struct Waiter {
    std::mutex m;
    std::condition_variable cv;
    bool done = false;
    int status = GAIL_OK;
    gail_packet pkt{};
};

struct Item {
    long long index = 0;
    uint8_t* buf = nullptr;  // from the encoder's buffer pool
    void* opaque = nullptr;
    Waiter* waiter = nullptr;
};

struct Done {
    std::vector<uint8_t> cs;
    void* opaque = nullptr;
    Waiter* waiter = nullptr;
    bool compliant = true;
    double mse = 0;
};

}  // namespace

struct gail_encoder {
    std::unique_ptr<StreamSpec> spec;
    gail_config cfg{};
    const GailHipPlugin* P = nullptr;
    GailHipDevice* dev = nullptr;
    std::vector<GailHipEncoder*> slots;
    std::string devices, error;
    std::atomic<bool> failed{false}, canceled{false};
    std::atomic<double> lastLambda{0.0};

    // input queue (bounded)
    std::mutex qm;
    std::condition_variable qcv;
    std::deque<Item> queue;
    size_t qcap = 4;
    long long nextIndex = 0;
    bool ended = false;
    // results
    std::mutex om;
    std::condition_variable ocv;
    std::map<long long, Done> pending;
    std::deque<gail_packet> out;
    long long nextOut = 0;
    int activeWorkers = 0;
    bool writerDone = false;
    std::vector<std::thread> workers;
    std::thread writer;
    // Host buffers frames are filled into, reused from frame to frame: pinned (the plugin's
    // host allocator) so that a worker's upload is a single DMA, and never zero-filled or
    // faulted in again. Every queued frame, every frame a worker holds and every send in
    // progress has one; they are freed with the encoder.
    std::mutex bm;
    std::vector<uint8_t*> freeBufs;
    std::vector<std::pair<uint8_t*, bool>> allBufs;  // buffer, pinned

    uint8_t* takeBuffer() {
        {
            std::lock_guard<std::mutex> l(bm);
            if (!freeBufs.empty()) {
                uint8_t* b = freeBufs.back();
                freeBufs.pop_back();
                return b;
            }
        }
        bool pinned = true;
        uint8_t* b = P->host_alloc ? (uint8_t*)P->host_alloc(spec->inBytes) : nullptr;
        if (!b) {
            pinned = false;
            b = (uint8_t*)std::malloc(std::max<size_t>(1, spec->inBytes));
        }
        if (!b) return nullptr;
        std::lock_guard<std::mutex> l(bm);
        allBufs.emplace_back(b, pinned);
        return b;
    }

    void giveBuffer(uint8_t* b) {
        if (!b) return;
        std::lock_guard<std::mutex> l(bm);
        freeBufs.push_back(b);
    }

    // Threads that copy parts of a frame beside the sending thread: a 4K frame is 51.8 MB,
    // more than one thread copies at the rate the device takes frames.
    std::unique_ptr<Pool> fillPool;

    void fillFrame(const gail_frame& f, uint8_t* dst) {
        const int helpers = fillPool ? fillPool->size() : 0;
        if (helpers == 0 || !spec->splitParts()) {
            spec->fill(f, dst);
            return;
        }
        const int parts = helpers + 1;
        std::mutex m;
        std::condition_variable cv;
        int left = helpers;
        for (int p = 1; p < parts; ++p)
            fillPool->post([&, p] {
                spec->fill(f, dst, p, parts);
                std::lock_guard<std::mutex> l(m);
                if (--left == 0) cv.notify_all();
            });
        spec->fill(f, dst, 0, parts);
        std::unique_lock<std::mutex> l(m);
        cv.wait(l, [&] { return left == 0; });
    }

    std::string errorText() {
        std::lock_guard<std::mutex> l(om);
        return error;
    }

    void log(gail_log_level lvl, const char* fmt, ...) {
        if (!cfg.log || lvl > cfg.log_level) return;
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        cfg.log(cfg.log_user, lvl, buf);
    }

    // Marks the stream failed (the first reason is kept) and fails every blocked encode call.
    void stop(const std::string& why) {
        {
            std::lock_guard<std::mutex> l(om);
            if (!failed.exchange(true)) error = why;
        }
        failWaiters(GAIL_ERROR_FAILED);
    }

    // Every frame still queued is dropped; a gail_encoder_encode caller waiting on one
    // returns `status`, as does one whose frame was coded but not yet delivered. Frames a
    // worker holds are failed by that worker when it hands its result over (deliverable()).
    // `failed` or `canceled` is set before this runs, so push() cannot queue a frame after it.
    void failWaiters(int status) {
        {
            std::lock_guard<std::mutex> l(qm);
            for (Item& it : queue)
                if (it.waiter) finishWaiter(it.waiter, status, nullptr);
            queue.clear();
        }
        qcv.notify_all();
        {
            std::lock_guard<std::mutex> l(om);
            for (auto& kv : pending)
                if (kv.second.waiter) finishWaiter(kv.second.waiter, status, nullptr);
            pending.clear();
        }
        ocv.notify_all();
    }

    int stoppedStatus() const { return canceled.load() ? GAIL_ERROR_CANCELLED : GAIL_ERROR_FAILED; }

    int push(const gail_frame* f, Waiter* w) {
        if (canceled.load()) return fail(GAIL_ERROR_CANCELLED, "canceled");
        if (failed.load()) return fail(GAIL_ERROR_FAILED, errorText());
        {
            std::lock_guard<std::mutex> l(qm);
            if (ended) return fail(GAIL_ERROR_STATE, "the stream has already ended");
        }
        std::string why;
        const int st = spec->checkFrame(*f, why);
        if (st != GAIL_OK) return fail(st, why);
        Item it;
        it.buf = takeBuffer();
        if (!it.buf) return fail(GAIL_ERROR_NOMEM, "no host memory for the frame");
        fillFrame(*f, it.buf);
        it.opaque = f->opaque;
        it.waiter = w;
        std::unique_lock<std::mutex> l(qm);
        qcv.wait(l, [&] { return queue.size() < qcap || failed.load() || canceled.load() || ended; });
        auto refuse = [&](int status, const std::string& why) {
            giveBuffer(it.buf);
            return fail(status, why);
        };
        if (canceled.load()) return refuse(GAIL_ERROR_CANCELLED, "canceled");
        if (failed.load()) return refuse(GAIL_ERROR_FAILED, errorText());
        // after send(NULL) or finish the workers may have gone: nothing queued now would run
        if (ended) return refuse(GAIL_ERROR_STATE, "the stream has already ended");
        it.index = nextIndex++;
        queue.push_back(std::move(it));
        qcv.notify_all();
        return GAIL_OK;
    }

    void workerLoop(GailHipEncoder* slot) {
        const std::string tag = "hip";
        for (;;) {
            Item it;
            {
                std::unique_lock<std::mutex> l(qm);
                qcv.wait(l, [&] { return !queue.empty() || ended || failed.load() || canceled.load(); });
                if (failed.load() || canceled.load() || queue.empty()) break;
                it = std::move(queue.front());
                queue.pop_front();
                qcv.notify_all();
            }
            Done d;
            d.opaque = it.opaque;
            d.waiter = it.waiter;
            try {
                const double thr = lastLambda.load() * spec->slopeFactor;
                P->encoder_set_slope(slot, thr);
                GailHipTier1 t1{};
                if (P->encoder_encode(slot, it.buf, &t1, nullptr, nullptr) != GAIL_HIP_OK)
                    throw std::runtime_error(P->last_error());
                // the device's rate control when the plugin ran it (the Tier-1 output then
                // holds each block's chosen passes and no more), else PCRD here
                auto rate = [&](const Tier1Ref& r) {
                    GailHipRate dr{};
                    if (P->encoder_rate && P->encoder_rate(slot, &dr) == GAIL_HIP_OK) {
                        RateResult x;
                        x.passes.resize(r.numBlocks);
                        for (size_t b = 0; b < r.numBlocks; ++b) x.passes[b] = (uint8_t)r.passes((int)b);
                        for (int c = 0; c < 3; ++c) x.lambdaComp[c] = dr.lambda[c];
                        x.distortion = dr.distortion;
                        return x;
                    }
                    return allocateRate(spec->L, r, spec->limits);
                };
                Tier1Ref ref{(const uint32_t*)t1.info, t1.passLen, t1.passDist, t1.data, (size_t)t1.numInfo};
                RateResult rr = rate(ref);
                double lam = std::min({rr.lambdaComp[0], rr.lambdaComp[1], rr.lambdaComp[2]});
                if (thr > 0 && lam < thr) {
                    // an easier frame than the last: passes the rate control would keep may
                    // have been skipped; code it again without early termination
                    P->encoder_set_slope(slot, 0.0);
                    if (P->encoder_encode(slot, it.buf, &t1, nullptr, nullptr) != GAIL_HIP_OK)
                        throw std::runtime_error(P->last_error());
                    ref = Tier1Ref{(const uint32_t*)t1.info, t1.passLen, t1.passDist, t1.data, (size_t)t1.numInfo};
                    rr = rate(ref);
                    lam = std::min({rr.lambdaComp[0], rr.lambdaComp[1], rr.lambdaComp[2]});
                }
                buildCodestream(spec->L, ref, rr.passes.data(), &d.cs, nullptr, tag, true);
                lastLambda.store(lam);
                d.mse = rr.distortion / ((double)spec->width * spec->height * 3);
                if (cfg.check) {
                    const std::string why = checkDciCodestream(d.cs.data(), d.cs.size(), spec->eyeFps(), spec->fourk());
                    d.compliant = why.empty();
                    if (!d.compliant) log(GAIL_LOG_WARNING, "frame %lld failed the DCI check: %s", it.index, why.c_str());
                }
            } catch (const std::exception& e) {
                giveBuffer(it.buf);
                stop(e.what());
                if (it.waiter) finishWaiter(it.waiter, GAIL_ERROR_FAILED, nullptr);
                break;
            }
            giveBuffer(it.buf);
            {
                std::lock_guard<std::mutex> l(om);
                if (!deliverable()) {
                    // canceled or failed while this frame was on the device: the writer
                    // delivers nothing more, so its waiter is answered here
                    if (d.waiter) finishWaiter(d.waiter, stoppedStatus(), nullptr);
                    continue;
                }
                pending.emplace(it.index, std::move(d));
            }
            ocv.notify_all();
        }
        {
            std::lock_guard<std::mutex> l(om);
            --activeWorkers;
        }
        ocv.notify_all();
    }

    // With om held: whether the writer will still take a frame put into `pending`.
    bool deliverable() const { return !writerDone && !failed.load() && !canceled.load(); }

    static void finishWaiter(Waiter* w, int status, const gail_packet* p) {
        std::lock_guard<std::mutex> l(w->m);
        w->status = status;
        if (p) w->pkt = *p;
        w->done = true;
        w->cv.notify_all();
    }

    void writerLoop() {
        for (;;) {
            Done d;
            long long idx;
            {
                std::unique_lock<std::mutex> l(om);
                ocv.wait(l, [&] { return pending.count(nextOut) || activeWorkers == 0 || failed.load() || canceled.load(); });
                if (failed.load() || canceled.load()) break;
                auto it = pending.find(nextOut);
                if (it == pending.end()) break;
                d = std::move(it->second);
                pending.erase(it);
                idx = nextOut++;
            }
            gail_packet p{};
            p.size = d.cs.size();
            p.index = idx;
            p.opaque = d.opaque;
            p.device = 0;
            p.compliant = d.compliant ? 1 : 0;
            p.psnr_estimate = d.mse > 0 ? 10 * std::log10(4095.0 * 4095.0 / d.mse) : 99.0;
            if (!d.waiter && cfg.on_packet) {
                p.data = d.cs.data();
                cfg.on_packet(cfg.user, &p);
                continue;
            }
            p.data = (uint8_t*)std::malloc(std::max<size_t>(1, d.cs.size()));
            if (!p.data) {
                stop("out of host memory");
                break;
            }
            std::memcpy(p.data, d.cs.data(), d.cs.size());
            if (d.waiter) {
                finishWaiter(d.waiter, GAIL_OK, &p);
                continue;
            }
            std::lock_guard<std::mutex> l(om);
            out.push_back(p);
            ocv.notify_all();
        }
        // wake blocked waiters of frames that will not be delivered; once writerDone is set
        // (under om) no worker adds to `pending` (deliverable())
        {
            std::lock_guard<std::mutex> l(om);
            for (auto& kv : pending)
                if (kv.second.waiter) finishWaiter(kv.second.waiter, stoppedStatus(), nullptr);
            pending.clear();
            writerDone = true;
        }
        ocv.notify_all();
        // a frame queued behind a failure that no worker will take
        if (failed.load() || canceled.load()) failWaiters(stoppedStatus());
    }

    void shutdown() {
        {
            std::lock_guard<std::mutex> l(qm);
            ended = true;
        }
        qcv.notify_all();
        for (auto& t : workers)
            if (t.joinable()) t.join();
        ocv.notify_all();
        if (writer.joinable()) writer.join();
        fillPool.reset();
        for (auto* s : slots) P->encoder_destroy(s);
        slots.clear();
        if (dev) P->device_destroy(dev);
        dev = nullptr;
        {
            std::lock_guard<std::mutex> lb(bm);
            for (auto& b : allBufs) {
                if (b.second) P->host_free(b.first);
                else std::free(b.first);
            }
            allBufs.clear();
            freeBufs.clear();
        }
        std::lock_guard<std::mutex> l(om);
        for (auto& p : out) std::free(p.data);
        out.clear();
    }
};

extern "C" {

GAIL_API const char* gail_last_error(void) { return lastErrorText().c_str(); }

GAIL_API const char* gail_status_string(int s) {
    switch (s) {
    case GAIL_OK: return "ok";
    case GAIL_EAGAIN: return "try again";
    case GAIL_EOF: return "end of stream";
    case GAIL_ERROR_INVALID: return "invalid argument";
    case GAIL_ERROR_UNSUPPORTED: return "unsupported";
    case GAIL_ERROR_NOMEM: return "out of memory";
    case GAIL_ERROR_FAILED: return "failed";
    case GAIL_ERROR_STATE: return "not allowed now";
    case GAIL_ERROR_CANCELLED: return "cancelled";
    default: return "unknown status";
    }
}

GAIL_API unsigned gail_version(void) { return GAIL_VERSION; }
GAIL_API const char* gail_version_string(void) { return "1.1.0 (hip; gail_hip)"; }

GAIL_API int gail_container_size(gail_profile p, gail_container c, int* w, int* h) {
    if ((p != GAIL_PROFILE_2K && p != GAIL_PROFILE_4K) || c < GAIL_CONTAINER_FLAT || c > GAIL_CONTAINER_FULL)
        return fail(GAIL_ERROR_INVALID, "invalid profile or container");
    int W, H;
    containerSize(p == GAIL_PROFILE_4K, (int)c, W, H);
    if (w) *w = W;
    if (h) *h = H;
    return GAIL_OK;
}

GAIL_API void gail_color_preset(gail_preset preset, gail_color* c) {
    std::memset(c, 0, sizeof *c);
    c->mode = GAIL_COLOR_MANAGED;
    c->transfer = GAIL_TRANSFER_BT1886;
    c->primaries = GAIL_PRIMARIES_BT709;
    c->tone_map = GAIL_TONE_MAP_SIGNAL;
    c->tone_norm = GAIL_TONE_NORM_CHANNEL;
    c->sdr_white = 48.0;
    c->hdr_peak = 1000.0;
    c->hdr_ref_white = 203.0;
    c->cinema_ref_white = 16.0;
    c->cinema_peak = 48.0;
    c->knee = 0.25;
    c->gamut = GAIL_GAMUT_NONE;
    c->gamut_threshold = 0.8;
    switch (preset) {
    case GAIL_PRESET_SDR_BT2020: c->primaries = GAIL_PRIMARIES_BT2020; break;
    case GAIL_PRESET_HDR_PQ_BT2020:
        c->transfer = GAIL_TRANSFER_PQ;
        c->primaries = GAIL_PRIMARIES_BT2020;
        break;
    case GAIL_PRESET_HDR_HLG_BT2020:
        c->transfer = GAIL_TRANSFER_HLG;
        c->primaries = GAIL_PRIMARIES_BT2020;
        break;
    default: break;
    }
}

GAIL_API void gail_config_init(gail_config* c, size_t size) {
    if (!c || size < sizeof(size_t)) return;
    gail_config d;
    std::memset(&d, 0, sizeof d);
    d.struct_size = size;  // the caller's header: fields past this library's end stay zero
    d.profile = GAIL_PROFILE_2K;
    d.container = GAIL_CONTAINER_FLAT;
    d.fps = 24.0;
    d.max_mbps = 230.0;
    d.input.format = GAIL_PIX_RGB48;
    gail_color_preset(GAIL_PRESET_SDR_BT709, &d.color);
    d.backend = GAIL_BACKEND_AUTO;
    d.check = 1;
    d.log_level = GAIL_LOG_WARNING;
    std::memset(c, 0, size);
    std::memcpy(c, &d, std::min(size, sizeof d));
}

GAIL_API void gail_packet_free(gail_packet* p) {
    if (!p) return;
    std::free(p->data);
    p->data = nullptr;
    p->size = 0;
}

GAIL_API int gail_encoder_create(const gail_config* c, gail_encoder** encoder) {
    if (!encoder) return fail(GAIL_ERROR_INVALID, "encoder pointer is NULL");
    *encoder = nullptr;
    // gail.h's compatibility rules: a config at least as large as this header's, and any
    // fields of a newer header left at zero
    if (!c || c->struct_size < sizeof(gail_config))
        return fail(GAIL_ERROR_INVALID, "config not initialized with gail_config_defaults()");
    if (c->struct_size > sizeof(gail_config)) {
        const uint8_t* extra = reinterpret_cast<const uint8_t*>(c) + sizeof(gail_config);
        for (size_t i = 0; i < c->struct_size - sizeof(gail_config); ++i)
            if (extra[i])
                return fail(GAIL_ERROR_UNSUPPORTED,
                            std::string("the config uses settings this libgail (") + gail_version_string() + ") lacks");
    }
    auto e = std::make_unique<gail_encoder>();
    try {
        e->spec = StreamSpec::make(*c);
        e->cfg = e->spec->cfg;
        gail_config& cfg = e->cfg;
        const PluginHandle& ph = loadPlugin();
        if (!ph.table) throw ApiError(GAIL_ERROR_UNSUPPORTED, ph.error);
        e->P = ph.table;
        const int n = e->P->device_count();
        if (n < 0) throw ApiError(GAIL_ERROR_UNSUPPORTED, std::string("HIP runtime unavailable: ") + e->P->last_error());
        if (n == 0) throw ApiError(GAIL_ERROR_UNSUPPORTED, "no HIP device");
        const int index = (cfg.devices && cfg.num_devices > 0) ? cfg.devices[0] : 0;
        if (index < 0 || index >= n) throw ApiError(GAIL_ERROR_INVALID, "no such device");
        if (e->P->device_usable(index) != GAIL_HIP_OK) throw ApiError(GAIL_ERROR_UNSUPPORTED, e->P->last_error());
        if (e->P->device_create(index, nullptr, nullptr, &e->dev) != GAIL_HIP_OK)
            throw ApiError(GAIL_ERROR_FAILED, e->P->last_error());
        e->devices = e->P->device_name(e->dev);
        // rate control on the device, as the served form runs it (byte-identical to PCRD on
        // the host): a worker then only writes the packets of the passes the device kept
        GailHipEncodeSettings es = e->spec->settings;
        es.rcMaxTotal = e->spec->limits.maxTotal;
        es.rcMaxComp = e->spec->limits.maxComp;
        // frames in flight: as asked (default 2), fewer when device memory is short
        const uint64_t per = e->P->encoder_bytes(e->dev, &e->spec->plainLayout, &es);
        GailHipMemory mem{};
        const bool known = e->P->device_memory(e->dev, &mem) == GAIL_HIP_OK && mem.budgeted;
        const int want = fitFrames(known ? mem.available : 0, mem.total, per,
                                   cfg.frames_in_flight > 0 ? cfg.frames_in_flight : 2, cfg.max_device_memory);
        for (int i = 0; i < want; ++i) {
            GailHipEncoder* slot = nullptr;
            const int st = e->P->encoder_create(e->dev, &e->spec->plainLayout, &es, &slot);
            if (st != GAIL_HIP_OK) {
                if (e->slots.empty())
                    throw ApiError(st == GAIL_HIP_DEVICE_MEMORY ? GAIL_ERROR_NOMEM : GAIL_ERROR_FAILED, e->P->last_error());
                e->log(GAIL_LOG_WARNING, "frames in flight reduced to %d: %s", (int)e->slots.size(), e->P->last_error());
                break;
            }
            e->slots.push_back(slot);
        }
        e->qcap = e->slots.size() + 2;
        e->fillPool.reset(new Pool(3));
        e->activeWorkers = (int)e->slots.size();
        for (auto* slot : e->slots) e->workers.emplace_back([ep = e.get(), slot] { ep->workerLoop(slot); });
        e->writer = std::thread([ep = e.get()] { ep->writerLoop(); });
        *encoder = e.release();
        return GAIL_OK;
    } catch (const ApiError& x) {
        e->shutdown();
        return fail(x.status, x.what());
    } catch (const std::exception& x) {
        e->shutdown();
        return fail(GAIL_ERROR_FAILED, x.what());
    }
}

GAIL_API void gail_encoder_destroy(gail_encoder* e) {
    if (!e) return;
    e->shutdown();
    delete e;
}

GAIL_API void gail_encoder_frame_size(const gail_encoder* e, int* w, int* h) {
    if (w) *w = e && e->spec ? e->spec->width : 0;
    if (h) *h = e && e->spec ? e->spec->height : 0;
}

GAIL_API const char* gail_encoder_devices(const gail_encoder* e) { return e ? e->devices.c_str() : ""; }

GAIL_API int gail_encoder_send(gail_encoder* e, const gail_frame* f) {
    if (!e) return fail(GAIL_ERROR_INVALID, "null encoder");
    if (e->canceled.load()) return fail(GAIL_ERROR_CANCELLED, "canceled");
    if (e->failed.load()) return fail(GAIL_ERROR_FAILED, e->errorText());
    if (!f) {
        {
            std::lock_guard<std::mutex> l(e->qm);
            e->ended = true;
        }
        e->qcv.notify_all();
        return GAIL_OK;
    }
    return e->push(f, nullptr);
}

GAIL_API int gail_encoder_receive(gail_encoder* e, gail_packet* pkt, int wait) {
    if (!e || !pkt) return fail(GAIL_ERROR_INVALID, "null argument");
    if (e->cfg.on_packet) return fail(GAIL_ERROR_STATE, "packets go to on_packet");
    std::unique_lock<std::mutex> l(e->om);
    for (;;) {
        // after a cancel nothing more is handed out; after a failure, what was coded is
        if (e->canceled.load()) return fail(GAIL_ERROR_CANCELLED, "canceled");
        if (!e->out.empty()) {
            *pkt = e->out.front();
            e->out.pop_front();
            return GAIL_OK;
        }
        if (e->failed.load()) return fail(GAIL_ERROR_FAILED, e->error);
        if (e->writerDone) return GAIL_EOF;
        if (!wait) return GAIL_EAGAIN;
        e->ocv.wait(l);
    }
}

GAIL_API int gail_encoder_encode(gail_encoder* e, const gail_frame* f, gail_packet* pkt) {
    if (!e || !f || !pkt) return fail(GAIL_ERROR_INVALID, "null argument");
    Waiter w;
    const int r = e->push(f, &w);
    if (r != GAIL_OK) return r;
    int status;
    {
        std::unique_lock<std::mutex> l(w.m);
        w.cv.wait(l, [&] { return w.done; });
        status = w.status;
    }
    // w.m is released before om is taken (errorText): the writer and workers take them the
    // other way round
    if (status != GAIL_OK) return fail(status, status == GAIL_ERROR_CANCELLED ? "canceled" : e->errorText());
    *pkt = w.pkt;
    return GAIL_OK;
}

GAIL_API int gail_encoder_finish(gail_encoder* e) {
    if (!e) return fail(GAIL_ERROR_INVALID, "null encoder");
    {
        std::lock_guard<std::mutex> l(e->qm);
        e->ended = true;
    }
    e->qcv.notify_all();
    std::unique_lock<std::mutex> l(e->om);
    e->ocv.wait(l, [&] { return e->writerDone; });
    if (e->canceled.load()) return fail(GAIL_ERROR_CANCELLED, "canceled");
    if (e->failed.load()) return fail(GAIL_ERROR_FAILED, e->error);
    return GAIL_OK;
}

GAIL_API void gail_encoder_cancel(gail_encoder* e) {
    if (!e || e->canceled.exchange(true)) return;
    // blocked sends, encodes and receives return CANCELLED (gail.h:354-358)
    e->failWaiters(GAIL_ERROR_CANCELLED);
}

GAIL_API const char* gail_encoder_error(const gail_encoder* e) {
    return e && e->failed.load() ? e->error.c_str() : nullptr;
}

GAIL_API int gail_enumerate_devices(gail_backend backend, gail_device_info* out, int max) {
    if (backend != GAIL_BACKEND_AUTO && backend != GAIL_BACKEND_HIP) return 0;
    const PluginHandle& ph = loadPlugin();
    if (!ph.table) return 0;
    const int n = ph.table->device_count();
    for (int i = 0; i < n && i < max; ++i) {
        GailHipDeviceDesc d{};
        if (ph.table->device_desc(i, &d) != GAIL_HIP_OK) continue;
        gail_device_info& o = out[i];
        std::memset(&o, 0, sizeof o);
        o.backend = GAIL_BACKEND_HIP;
        o.index = i;
        std::snprintf(o.name, sizeof o.name, "%s", d.name);
        std::snprintf(o.pci_bus_id, sizeof o.pci_bus_id, "%s", d.pciBusId);
        o.memory_bytes = d.memoryBytes;
    }
    return std::max(n, 0);
}

}  // extern "C"
