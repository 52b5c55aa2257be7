// The plugin table of libgail_hip.so (include/gail_hip_plugin.h): plain-C entry points over
// the device engine, with exceptions turned into status codes and a thread-local message
// (the pattern of the reference's cuda_plugin.cpp:48-70, 220-249).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <new>
#include <stdexcept>
#include <string>

#include "decoder.h"
#include "engine.h"
#include "errors.h"
#include "gail_hip_plugin.h"
#include "host_pool.h"

using namespace gail_amd;

struct GailHipDevice {
    int index = 0;
    std::string name;
    GailHipLog log = nullptr;
    void* user = nullptr;
};

struct GailHipEncoder {
    std::unique_ptr<Job> job;
    std::vector<uint16_t> dbgCodes;
    std::vector<float> dbgWavelet;
    std::vector<uint32_t> dbgCoeffs;
};

struct GailHipDecoder {
    std::unique_ptr<Decoder> dec;
};

struct GailHipBatch {
    std::unique_ptr<Job> job;
    std::vector<uint16_t> dbgCodes;
    std::vector<float> dbgWavelet;
    std::vector<uint32_t> dbgCoeffs;
};

namespace {

thread_local std::string g_error;

template <class F> int guarded(F&& f) {
    try {
        f();
        return GAIL_HIP_OK;
    } catch (const DeviceOutOfMemory& e) {
        g_error = e.what();
        return GAIL_HIP_DEVICE_MEMORY;
    } catch (const std::bad_alloc&) {
        g_error = "out of host memory";
        return GAIL_HIP_HOST_MEMORY;
    } catch (const std::logic_error& e) {
        g_error = e.what();
        return GAIL_HIP_MISUSE;
    } catch (const std::exception& e) {
        g_error = e.what();
        return GAIL_HIP_ERROR;
    } catch (...) {
        g_error = "unknown error";
        return GAIL_HIP_ERROR;
    }
}

// GailHipLayout -> Layout: the geometry is rebuilt from the size and levels, the subbands
// (which carry the step sizes) are taken as they are (cuda_plugin.h:245-253).
// Geometry of a size and decomposition, built once per process (the served form binds the
// same few DCI sizes over and over).
const Layout& cachedGeometry(int width, int height, bool fourk, int levels, int guardBits) {
    static std::mutex m;
    static std::map<std::tuple<int, int, bool, int, int>, std::unique_ptr<Layout>> cache;
    std::lock_guard<std::mutex> l(m);
    auto& p = cache[std::make_tuple(width, height, fourk, levels, guardBits)];
    if (!p) p = std::make_unique<Layout>(makeGeometry(width, height, fourk, levels, guardBits));
    return *p;
}

Layout toLayout(const GailHipLayout& p, int maxMb = 22) {
    if (!p.bands || p.numBands <= 0) throw std::logic_error("frame layout is empty");
    if (p.width <= 0 || p.height <= 0 || p.width > 65535 || p.height > 65535 || p.levels < 1 || p.levels > 8)
        throw std::logic_error("frame layout size or levels out of range");
    Layout L = cachedGeometry(p.width, p.height, p.profile == 1, p.levels, p.guardBits);
    if ((int)L.bands.size() != p.numBands || L.numCblks() != p.numCblks)
        throw std::logic_error("frame layout does not have the geometry of its size and levels");
    L.baseStep = p.baseStep;
    for (int i = 0; i < p.numBands; ++i) {
        const GailHipSubband& s = p.bands[i];
        Subband& b = L.bands[i];
        if (s.comp != b.comp || s.res != b.res || s.orient != b.orient || s.w != b.w || s.h != b.h)
            throw std::logic_error("subband does not match the frame geometry");
        if (s.Mb < 0 || s.Mb > maxMb)
            throw std::logic_error("subband with more than " + std::to_string(maxMb) + " magnitude bit-planes");
        b.eps = s.eps;
        b.mu = s.mu;
        b.Mb = s.Mb;
        b.delta = s.delta;
        b.weight = s.weight;
    }
    return L;
}

FrameConfig toFrame(const GailHipLayout& l, const GailHipEncodeSettings& s) {
    FrameConfig f;
    f.layout = toLayout(l);
    const GailHipFrameFormat& in = s.input;
    if (in.kind < 0 || in.kind > 2) throw std::logic_error("unknown input layout");
    f.input.kind = in.kind;
    f.input.width = in.width ? in.width : l.width;
    f.input.height = in.height ? in.height : l.height;
    f.input.aw = in.activeW;
    f.input.ah = in.activeH;
    f.input.ox = in.ox;
    f.input.oy = in.oy;
    f.input.sx = in.sx;
    f.input.sy = in.sy;
    f.input.bits = in.bits;
    f.input.rgb = in.rgb;
    f.input.fullRange = in.fullRange;
    f.input.chroma = in.chroma;
    f.input.kr = in.kr;
    f.input.kb = in.kb;
    if (in.kind == 1) {
        if (in.bits < 8 || in.bits > 16) throw std::logic_error("planar input bit depth out of range");
        if (in.activeW <= 0 || in.activeH <= 0 || in.ox < 0 || in.oy < 0 || in.ox + in.activeW > f.input.width ||
            in.oy + in.activeH > f.input.height)
            throw std::logic_error("planar picture does not fit the container");
    }
    f.rcMaxTotal = s.rcMaxTotal;
    f.rcMaxComp = s.rcMaxComp;
    const GailHipColor& c = s.color;
    if (c.transfer < 0 || c.transfer > 5 || c.primaries < 0 || c.primaries > 4 || c.toneMap < 0 || c.toneMap > 3 ||
        c.toneNorm < 0 || c.toneNorm > 2 || c.gamut < 0 || c.gamut > 2)
        throw std::logic_error("colour enumeration out of range");
    if (c.cinemaPeak > 52.37 || !(c.gamutThreshold >= 0 && c.gamutThreshold < 1) || (c.knee != 0 && !(c.knee > 0 && c.knee < 1)))
        throw std::logic_error("colour parameter out of range");
    f.color.transfer = c.transfer;
    f.color.primaries = c.primaries;
    f.color.toneMap = c.toneMap;
    f.color.toneNorm = c.toneNorm;
    f.color.gamut = c.gamut;
    f.color.sdrWhite = c.sdrWhite > 0 ? c.sdrWhite : 48.0;
    f.color.hdrPeak = c.hdrPeak > 0 ? c.hdrPeak : 1000.0;
    f.color.cinemaPeak = c.cinemaPeak > 0 ? c.cinemaPeak : 48.0;
    f.color.hdrRefWhite = c.hdrRefWhite > 0 ? c.hdrRefWhite : 203.0;
    f.color.cinemaRefWhite = c.cinemaRefWhite > 0 ? c.cinemaRefWhite : 16.0;
    f.color.knee = c.knee > 0 ? c.knee : 0.25;
    f.color.gamutThreshold = c.gamutThreshold;
    if (c.lut3dData && c.lut3dCount) {
        if (c.lut3dSize < 2 || c.lut3dSize > 256 || c.lut3dCount != (uint64_t)c.lut3dSize * c.lut3dSize * c.lut3dSize * 3)
            throw std::logic_error("3D LUT data does not match its size");
        f.color.lut3dSize = c.lut3dSize;
        f.color.lut3d = std::make_shared<const std::vector<float>>(c.lut3dData, c.lut3dData + c.lut3dCount);
    } else if (c.lut3d && c.lut3d[0]) {
        int n = 0;
        f.color.lut3d = std::make_shared<const std::vector<float>>(readCubeLut(c.lut3d, n));
        f.color.lut3dSize = n;
    }
    return f;
}

JobOptions toOptions(const GailHipBatchOptions* o) {
    JobOptions j;
    if (!o) return j;
    if (o->dwtTile >= 0 && o->dwtTile < kDwtShapes) j.dwtTile = o->dwtTile;
    if (o->cmWaves == 1 || o->cmWaves == 2 || o->cmWaves == 4 || o->cmWaves == 8) j.cmWaves = o->cmWaves;
    if (o->mqGroup == 64 || o->mqGroup == 128 || o->mqGroup == 256) j.mqGroup = o->mqGroup;
    if (o->mqRead == 16 || o->mqRead == 64) j.mqRead = o->mqRead;
    j.useGraph = o->useGraph != 0;
    if (o->poolBytesPerCoef > 0) j.poolBytesPerCoef = o->poolBytesPerCoef;
    return j;
}

void fillTier1(const Job& job, int i, GailHipTier1* out) {
    const Tier1View v = job.result(i);
    out->info = reinterpret_cast<const GailHipCblkInfo*>(v.info5);
    out->passLen = v.passLen;
    out->passDist = v.passDist;
    out->data = v.data;
    out->numInfo = v.numBlocks;
    out->numPasses = v.numPasses;
    out->numData = v.numData;
}

void fillTimings(const StageTimes& t, GailHipFrameTimings* out) {
    std::memset(out, 0, sizeof *out);
    out->upload = t.upload;
    out->color = t.color;
    out->dwt = t.dwt;
    out->t1cm = t.t1cm;
    out->t1mq = t.t1mq + t.compact;
    out->t1 = t.t1cm + t.t1mq + t.compact;
    out->download = t.download;
    out->gpu = t.gpu;
    out->total = t.gpu;
    out->codedBytes = t.bytes;
}

template <class T> void debugTaps(T* e, int i, GailHipDebug* dbg) {
    e->job->debugCodes(i, e->dbgCodes);
    dbg->xyz12 = e->dbgCodes.data();
    dbg->numXyz12 = e->dbgCodes.size();
    if (dbg->full) {
        e->job->debugWavelet(i, e->dbgWavelet);
        e->job->debugCoeffs(i, e->dbgCoeffs);
        dbg->wavelet = e->dbgWavelet.data();
        dbg->numWavelet = e->dbgWavelet.size();
        dbg->coeffs = e->dbgCoeffs.data();
        dbg->numCoeffs = e->dbgCoeffs.size();
    }
}

// ---- table entries
const char* lastError() { return g_error.c_str(); }

int deviceCount() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return n;
}

int deviceDesc(int index, GailHipDeviceDesc* out) {
    return guarded([&] {
        hipDeviceProp_t p;
        hipCheck(hipGetDeviceProperties(&p, index), "hipGetDeviceProperties");
        std::memset(out, 0, sizeof *out);
        out->index = index;
        out->computeUnits = p.multiProcessorCount;
        out->memoryBytes = p.totalGlobalMem;
        std::snprintf(out->name, sizeof out->name, "%s", p.name);
        std::snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
        std::snprintf(out->pciBusId, sizeof out->pciBusId, "%04x:%02x:%02x.0", p.pciDomainID, p.pciBusID, p.pciDeviceID);
    });
}

int deviceUsable(int index) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, index) != hipSuccess) {
        (void)hipGetLastError();
        g_error = "no such device";
        return GAIL_HIP_ERROR;
    }
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        g_error = std::string("no code object for ") + p.gcnArchName + " (built for gfx950)";
        return GAIL_HIP_NO_CODE;
    }
    return guarded([&] { DeviceContext::get(index); });
}

int deviceCreate(int index, GailHipLog log, void* user, GailHipDevice** out) {
    return guarded([&] {
        DeviceContext& c = DeviceContext::get(index);
        auto d = std::make_unique<GailHipDevice>();
        d->index = index;
        d->name = c.name() + " (HIP " + c.arch() + ")";
        d->log = log;
        d->user = user;
        if (log) log(user, 2, ("hip: device " + std::to_string(index) + " " + d->name).c_str());
        *out = d.release();
    });
}

void deviceDestroy(GailHipDevice* d) { delete d; }
const char* deviceName(GailHipDevice* d) { return d->name.c_str(); }

int deviceMemory(GailHipDevice* d, GailHipMemory* out) {
    return guarded([&] {
        size_t f = 0, t = 0;
        hipCheck(hipSetDevice(d->index), "hipSetDevice");
        hipCheck(hipMemGetInfo(&f, &t), "hipMemGetInfo");
        out->total = t;
        out->available = f;
        out->budgeted = 1;
    });
}

uint64_t encoderBytes(GailHipDevice*, const GailHipLayout* l, const GailHipEncodeSettings* s) {
    try {
        return Job::estimateBytes({toFrame(*l, *s)}, JobOptions{});
    } catch (...) {
        return 0;
    }
}

int encoderCreate(GailHipDevice* d, const GailHipLayout* l, const GailHipEncodeSettings* s, GailHipEncoder** out) {
    return guarded([&] {
        auto e = std::make_unique<GailHipEncoder>();
        e->job = std::make_unique<Job>(d->index, std::vector<FrameConfig>{toFrame(*l, *s)}, JobOptions{});
        *out = e.release();
    });
}

void encoderDestroy(GailHipEncoder* e) { delete e; }
void* encoderInput(GailHipEncoder* e) {
    void* p = nullptr;
    guarded([&] { p = e->job->hostInput(0); });
    return p;
}
void encoderSetSlope(GailHipEncoder* e, double t) { e->job->setSlope(0, t); }

// Whether the device can copy from p by DMA: pinned host memory or device memory.
static bool dmaReadable(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is not an error here
        return false;
    }
    return a.type == hipMemoryTypeHost || a.type == hipMemoryTypeDevice;
}

// Copies a frame from pageable memory into pinned staging in eight pieces on the calling
// thread and three helpers: one thread's memcpy of a 4K frame (38.8 MB) takes about 3 ms.
static void stageFrame(void* dst, const void* src, size_t bytes) {
    static Pool* const helpers = new Pool(3);  // shared by every encoder; kept until exit
    constexpr int kPieces = 8;
    parallelChunks(*helpers, kPieces, [&](int k, int) {
        const size_t a = bytes * k / kPieces, b = bytes * (k + 1) / kPieces;
        std::memcpy((char*)dst + a, (const char*)src + a, b - a);
    });
}

int encoderEncode(GailHipEncoder* e, const void* frame, GailHipTier1* out, GailHipFrameTimings* timings,
                  GailHipDebug* debug) {
    return guarded([&] {
        if (!frame || !out) throw std::logic_error("encode needs a frame and an output");
        Job& j = *e->job;
        StageTimes t;
        if (frame == j.hostInput(0)) {
            j.encode(true, &t);
        } else if (dmaReadable(frame)) {  // one copy straight into the slot (not timed as upload)
            j.uploadFrom(0, frame);
            j.encode(false, &t);
        } else {  // pageable memory: through the slot's pinned staging buffer
            stageFrame(j.hostInput(0), frame, j.frame(0).input.bytes());
            j.encode(true, &t);
        }
        fillTier1(j, 0, out);
        if (timings) fillTimings(t, timings);
        if (debug) debugTaps(e, 0, debug);
    });
}

// The decoder carries up to 30 magnitude bit-planes per block (the block table says how many
// each has); the subbands' own Mb only bound them.
constexpr int kDecodeMaxMb = 31;

void checkDecodeSettings(const GailHipDecodeSettings* s, const GailHipLayout* l) {
    if (!s || !l) throw std::logic_error("null layout or decode settings");
    if (s->reduce < 0 || s->reduce > l->levels) throw std::runtime_error("more resolution levels left out than the stream has");
}

uint64_t decoderBytes(GailHipDevice*, const GailHipLayout* l, const GailHipDecodeSettings* s) {
    uint64_t r = 0;
    const int st = guarded([&] {
        checkDecodeSettings(s, l);
        r = Decoder::bytesFor(toLayout(*l, kDecodeMaxMb), s->reduce);
    });
    return st == GAIL_HIP_OK ? r : 0;
}

int decoderCreate(GailHipDevice* d, const GailHipLayout* l, const GailHipDecodeSettings* s, GailHipDecoder** out) {
    return guarded([&] {
        if (!d || !out) throw std::logic_error("null device or output");
        *out = nullptr;
        checkDecodeSettings(s, l);
        auto g = std::make_unique<GailHipDecoder>();
        g->dec = std::make_unique<Decoder>(d->index, toLayout(*l, kDecodeMaxMb), s->reduce, s->highBits != 0);
        *out = g.release();
    });
}

void decoderDestroy(GailHipDecoder* d) { delete d; }

int decoderDecode(GailHipDecoder* d, const GailHipLayout* frame, const uint8_t* data, size_t size,
                  const GailHipCblkCode* blocks, size_t numBlocks, void* dst, ptrdiff_t strideBytes,
                  GailHipDecodeTimings* t) {
    static_assert(sizeof(GailHipCblkCode) == sizeof(DecCode), "block record layout");
    return guarded([&] {
        if (!d || !frame) throw std::logic_error("null decoder or frame layout");
        const Layout L = toLayout(*frame, kDecodeMaxMb);
        DecodeTimes tm;
        const bool wave = t && t->wantWavelet;
        d->dec->decode(L, data, size, reinterpret_cast<const DecCode*>(blocks), numBlocks, dst, strideBytes, wave,
                       t ? &tm : nullptr);
        if (t) {
            t->t1 = tm.t1;
            t->dwt = tm.dwt;
            t->color = tm.color;
            t->total = tm.total;
            for (int c = 0; c < 3; ++c) t->wavelet[c] = wave ? d->dec->wavelet(c) : nullptr;
            t->numWavelet = wave ? d->dec->waveletSamples() : 0;
        }
    });
}

std::vector<FrameConfig> toFrames(const GailHipLayout* l, const GailHipEncodeSettings* s, int count) {
    if (count <= 0 || !l || !s) throw std::logic_error("a batch needs at least one frame");
    std::vector<FrameConfig> f;
    for (int i = 0; i < count; ++i) f.push_back(toFrame(l[i], s[i]));
    return f;
}

uint64_t batchBytes(GailHipDevice*, const GailHipLayout* l, const GailHipEncodeSettings* s, int count,
                    const GailHipBatchOptions* o) {
    try {
        return Job::estimateBytes(toFrames(l, s, count), toOptions(o));
    } catch (...) {
        return 0;
    }
}

int batchCreate(GailHipDevice* d, const GailHipLayout* l, const GailHipEncodeSettings* s, int count,
                const GailHipBatchOptions* o, GailHipBatch** out) {
    return guarded([&] {
        auto b = std::make_unique<GailHipBatch>();
        b->job = std::make_unique<Job>(d->index, toFrames(l, s, count), toOptions(o));
        *out = b.release();
    });
}

void batchDestroy(GailHipBatch* b) { delete b; }
void* batchInput(GailHipBatch* b, int i) {
    void* p = nullptr;
    guarded([&] { p = b->job->hostInput(i); });
    return p;
}
int batchUpload(GailHipBatch* b, int i, const void* frame) {
    return guarded([&] {
        if (i < 0 || i >= b->job->numFrames() || !frame) throw std::logic_error("no such frame slot");
        b->job->upload(i, frame);
    });
}
void batchSetSlope(GailHipBatch* b, int i, double t) {
    if (i >= 0 && i < b->job->numFrames()) b->job->setSlope(i, t);
}
int batchLaunch(GailHipBatch* b, int fromStaging) {
    return guarded([&] { b->job->launch(fromStaging != 0); });
}
int batchWait(GailHipBatch* b, GailHipTier1* outs, GailHipFrameTimings* timings) {
    return guarded([&] {
        StageTimes t;
        b->job->finish(&t);
        if (outs)
            for (int i = 0; i < b->job->numFrames(); ++i) fillTier1(*b->job, i, &outs[i]);
        if (timings) fillTimings(t, timings);
    });
}
int batchEncode(GailHipBatch* b, int fromStaging, GailHipTier1* outs, GailHipFrameTimings* timings) {
    const int r = batchLaunch(b, fromStaging);
    return r != GAIL_HIP_OK ? r : batchWait(b, outs, timings);
}
int batchDebug(GailHipBatch* b, int i, GailHipDebug* dbg) {
    return guarded([&] {
        if (i < 0 || i >= b->job->numFrames()) throw std::logic_error("no such frame slot");
        debugTaps(b, i, dbg);
    });
}

int batchStageMs(GailHipBatch* b, double* out, int n) {
    return guarded([&] {
        const StageTimes& t = b->job->lastTimes();
        const double v[12] = {t.upload, t.color, t.dwt, t.dwt0, t.t1cm, t.t1mq, t.compact, t.download, t.gpu,
                              t.symbols, t.bytes, (double)t.reruns};
        for (int i = 0; i < n && i < 12; ++i) out[i] = v[i];
    });
}

std::vector<SlotCapacity> toSlots(const GailHipSlot* s, int count) {
    if (count <= 0 || !s) throw std::logic_error("a batch needs at least one frame slot");
    std::vector<SlotCapacity> v;
    for (int i = 0; i < count; ++i) v.push_back(SlotCapacity{s[i].maxWidth, s[i].maxHeight, s[i].colour != 0});
    return v;
}

uint64_t slotsBytes(GailHipDevice*, const GailHipSlot* s, int count, const GailHipBatchOptions* o) {
    try {
        return Job::estimateBytes(toSlots(s, count), toOptions(o));
    } catch (...) {
        return 0;
    }
}

int slotsCreate(GailHipDevice* d, const GailHipSlot* s, int count, const GailHipBatchOptions* o, GailHipBatch** out) {
    return guarded([&] {
        auto b = std::make_unique<GailHipBatch>();
        b->job = std::make_unique<Job>(d->index, toSlots(s, count), toOptions(o));
        *out = b.release();
    });
}

int batchBind(GailHipBatch* b, const GailHipLayout* l, const GailHipEncodeSettings* s, int count) {
    return guarded([&] { b->job->bind(toFrames(l, s, count)); });
}

int batchQuery(GailHipBatch* b) {
    int r = 0;
    const int st = guarded([&] { r = b->job->done() ? 1 : 0; });
    return st == GAIL_HIP_OK ? r : -st;
}

int batchUploadFrom(GailHipBatch* b, int i, const void* src) {
    return guarded([&] {
        if (!src) throw std::logic_error("no source frame");
        b->job->uploadFrom(i, src);
    });
}

void* hostAllocEntry(uint64_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        g_error = "out of pinned host memory";
        return nullptr;
    }
    return p;
}

void hostFreeEntry(void* p) {
    if (p) hipHostFree(p);
}

int batchAutotune(GailHipBatch* b, int reps, GailHipBatchOptions* chosen, double* times, int ntimes) {
    return guarded([&] {
        TuneTimes t;
        const JobOptions o = b->job->autotune(reps, &t);
        if (chosen) {
            chosen->dwtTile = o.dwtTile;
            chosen->cmWaves = o.cmWaves;
            chosen->mqGroup = o.mqGroup;
            chosen->mqRead = o.mqRead;
            chosen->useGraph = o.useGraph ? 1 : 0;
            chosen->poolBytesPerCoef = o.poolBytesPerCoef;
        }
        // the four LDS tile shapes first and the strips last, as the table grew
        const double v[] = {t.dwt[0], t.dwt[1], t.dwt[2], t.dwt[3], t.cm[0],  t.cm[1], t.cm[2],
                            t.cm[3],  t.mq[0],  t.mq[1],  t.mq[2],  t.mq[3],  t.mq[4], t.mq[5],
                            t.dwt[4], t.dwt[5], t.dwt[6], t.dwt[7], t.dwt[8], t.dwt[9], t.dwt[10]};
        static_assert(sizeof(v) / sizeof(v[0]) == 10 + kDwtShapes, "every stage time once");
        const int nv = (int)(sizeof(v) / sizeof(v[0]));
        for (int i = 0; times && i < ntimes && i < nv; ++i) times[i] = v[i];
    });
}

void* deviceAllocEntry(GailHipDevice* d, uint64_t bytes) {
    void* p = nullptr;
    const int st = guarded([&] {
        hipCheck(hipSetDevice(d->index), "hipSetDevice");
        hipCheck(hipMalloc(&p, bytes ? bytes : 16), "device allocation");
    });
    return st == GAIL_HIP_OK ? p : nullptr;
}

void deviceFreeEntry(void* p) {
    if (p) hipFree(p);
}

int copyEntry(void* dst, const void* src, uint64_t bytes) {
    return guarded([&] { hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "copy"); });
}

int isDevicePointer(const void* p) {
    hipPointerAttribute_t a{};
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return a.type == hipMemoryTypeDevice ? 1 : 0;
}

int batchInputsReady(GailHipBatch* b) {
    return guarded([&] {
        if (!b) throw std::logic_error("null batch");
        b->job->inputsReady();
    });
}

void fillRate(const RcFrame* r, GailHipRate* out) {
    if (!r || !out) throw std::logic_error("the frame had no rate control on the device");
    for (int c = 0; c < 3; ++c) {
        out->lambda[c] = r->lambda[c];
        out->comp[c] = r->comp[c];
    }
    out->total = r->total;
    for (int k = 0; k < 6; ++k) out->tilePart[k] = r->tile_part[k];
    out->distortion = r->distortion;
    out->iterations = r->iterations;
    out->numTileParts = r->ntp;
}

int encoderRate(GailHipEncoder* e, GailHipRate* out) {
    return guarded([&] {
        if (!e) throw std::logic_error("null encoder");
        fillRate(e->job->rateResult(0), out);
    });
}

int batchRate(GailHipBatch* b, int i, GailHipRate* out) {
    return guarded([&] { fillRate(b->job->rateResult(i), out); });
}

#define GAIL_STR2(x) #x
#define GAIL_STR(x) GAIL_STR2(x)

const GailHipPlugin g_plugin = {
    GAIL_HIP_PLUGIN_ABI,
    (uint32_t)sizeof(GailHipPlugin),
    gail_hip_plugin_layout(),
    "gail_hip",
    "HIP " GAIL_STR(HIP_VERSION_MAJOR) "." GAIL_STR(HIP_VERSION_MINOR),
    "gfx950",
    lastError,
    deviceCount,
    deviceDesc,
    deviceUsable,
    deviceCreate,
    deviceDestroy,
    deviceName,
    deviceMemory,
    encoderBytes,
    encoderCreate,
    encoderDestroy,
    encoderInput,
    encoderSetSlope,
    encoderEncode,
    decoderBytes,
    decoderCreate,
    decoderDestroy,
    decoderDecode,
    batchBytes,
    batchCreate,
    batchDestroy,
    batchInput,
    batchUpload,
    batchSetSlope,
    batchEncode,
    batchLaunch,
    batchWait,
    batchDebug,
    batchStageMs,
    slotsBytes,
    slotsCreate,
    batchBind,
    batchQuery,
    batchUploadFrom,
    hostAllocEntry,
    hostFreeEntry,
    batchAutotune,
    deviceAllocEntry,
    deviceFreeEntry,
    copyEntry,
    isDevicePointer,
    batchRate,
    batchInputsReady,
    encoderRate,
};

}  // namespace

extern "C" const GailHipPlugin* gail_hip_plugin_get(uint32_t abi) { return abi == GAIL_HIP_PLUGIN_ABI ? &g_plugin : nullptr; }

extern "C" __attribute__((visibility("default"))) int gail_hip_decode_noalloc(
    int device, const GailHipLayout* frame, int reduce, int highBits, const uint8_t* data, size_t size,
    const GailHipCblkCode* blocks, size_t numBlocks, void* dst, ptrdiff_t strideBytes, void* stream) {
    return guarded([&] {
        if (!frame) throw std::logic_error("null frame layout");
        decodeNoAlloc(device, toLayout(*frame, kDecodeMaxMb), reduce, highBits != 0, data, size,
                      reinterpret_cast<const DecCode*>(blocks), numBlocks, dst, strideBytes, (hipStream_t)stream);
    });
}
