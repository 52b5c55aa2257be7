// Stand-in for libgail_hip.so, for the CPU tests of libgail.so's plugin loader and frame
// loop (the role of the reference's tests/test_cuda_plugin.cpp for its CUDA plugins). It
// exports the plugin entry point and table and needs no GPU: a device is a name, and an
// "encode" turns the frame into a tiny Tier-1 result on the CPU (one coding pass in every
// seventh code-block, whose bytes come from a hash of the input), so that libgail's rate
// control, Tier-2 and DCI check run on every frame and each frame's codestream depends on
// its pixels.
//
// STUB_MODE (compile time) picks what the loader sees:
//   0  a working plugin
//   1  a table of another ABI version, handed out whatever the caller asks for
//   2  the entry point turns the caller's ABI down (returns NULL)
//   3  a table whose size field is not sizeof(GailHipPlugin)
//   4  a table whose layout hash is not the library's
//   5  no entry point at all
//   6  a working table whose runtime finds no driver (device_count() < 0)
//
// A working stand-in also reads, at each call, what the test asks of it:
//   GAIL_STUB_FAIL_AT=n    the n-th frame coded on a device (1-based, by an encoder or a
//                          batch) and every later one fail as a lost device would
//   GAIL_STUB_SLOTS=k      encoder_create reports device memory exhausted after k encoders
//   GAIL_STUB_DELAY_MS=d   every encode, and every batch run, takes d ms longer
//   GAIL_STUB_MEM_MB=m     the device has m MiB (default 288 GiB), of which frame slots take
//                          what slots_bytes says; slots_create fails beyond it
//   GAIL_STUB_MEM_USED_MB=u  another process holds u MiB of it
// Each is read at every call; gail_stub_control() sets one from the test program.
// Batches (the served form's jobs) code on the CPU like the encoder: one frame slot holds
// any frame up to its size; slots_bytes charges 40 bytes per pixel of a slot.
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gail_hip_plugin.h"

#ifndef STUB_MODE
#define STUB_MODE 0
#endif

namespace {

thread_local std::string g_error = "no error";

long envLong(const char* name, long dflt) {
    // a value set through gail_stub_control() wins over the environment
    extern long controlValue(const char* name, long dflt);
    const char* v = std::getenv(name);
    return controlValue(name, v && *v ? std::strtol(v, nullptr, 10) : dflt);
}

// Knobs a test program sets while the plugin runs (getenv is not safe beside setenv).
struct Knob {
    const char* name;
    std::atomic<long> value;
    std::atomic<bool> set;
};
Knob g_knobs[] = {{"GAIL_STUB_FAIL_AT", {0}, {false}}, {"GAIL_STUB_SLOTS", {0}, {false}},
                  {"GAIL_STUB_DELAY_MS", {0}, {false}}, {"GAIL_STUB_MEM_MB", {0}, {false}},
                  {"GAIL_STUB_MEM_USED_MB", {0}, {false}}};

long controlValue(const char* name, long dflt) {
    for (Knob& k : g_knobs)
        if (!std::strcmp(k.name, name)) return k.set.load() ? k.value.load() : dflt;
    return dflt;
}

int failWith(int status, const char* msg) {
    g_error = msg;
    return status;
}

struct Device {
    std::string name = "Stand-in MI355X (HIP test plugin)";
    std::atomic<long> encodes{0};
    std::atomic<int> encoders{0};
    std::atomic<uint64_t> slotBytes{0};  // held by live batches
};

uint64_t deviceTotal() { return (uint64_t)envLong("GAIL_STUB_MEM_MB", 288 * 1024) << 20; }

uint64_t deviceAvailable(const Device* d) {
    const uint64_t used = d->slotBytes.load() + ((uint64_t)envLong("GAIL_STUB_MEM_USED_MB", 0) << 20);
    const uint64_t total = deviceTotal();
    return used < total ? total - used : 0;
}

struct FrameResult;
struct Encoder {
    Device* dev;
    int numCblks;
    size_t inBytes;
    double slope = 0;
    std::vector<GailHipCblkInfo> info;
    std::vector<uint32_t> passLen;
    std::vector<int32_t> passDist;
    std::vector<uint8_t> data;
};

const char* lastError() { return g_error.c_str(); }

int deviceCount() {
    if (STUB_MODE == 6) return failWith(-1, "no AMD GPU driver");
    return 1;
}

int deviceDesc(int index, GailHipDeviceDesc* out) {
    if (index != 0) return failWith(GAIL_HIP_MISUSE, "no such device");
    std::memset(out, 0, sizeof *out);
    out->index = 0;
    out->computeUnits = 256;
    out->memoryBytes = 288ull << 30;
    std::snprintf(out->name, sizeof out->name, "Stand-in MI355X");
    std::snprintf(out->arch, sizeof out->arch, "gfx950");
    std::snprintf(out->pciBusId, sizeof out->pciBusId, "0000:00:00.0");
    return GAIL_HIP_OK;
}

int deviceUsable(int index) { return index == 0 ? GAIL_HIP_OK : failWith(GAIL_HIP_MISUSE, "no such device"); }

int deviceCreate(int index, GailHipLog log, void* user, GailHipDevice** out) {
    if (index != 0) return failWith(GAIL_HIP_MISUSE, "no such device");
    Device* d = new Device();
    if (log) log(user, 2, "stub: device 0");
    *out = reinterpret_cast<GailHipDevice*>(d);
    return GAIL_HIP_OK;
}

void deviceDestroy(GailHipDevice* d) { delete reinterpret_cast<Device*>(d); }
const char* deviceName(GailHipDevice* d) { return reinterpret_cast<Device*>(d)->name.c_str(); }

int deviceMemory(GailHipDevice* d, GailHipMemory* out) {
    std::memset(out, 0, sizeof *out);
    out->total = deviceTotal();
    out->available = deviceAvailable(reinterpret_cast<Device*>(d));
    out->budgeted = 1;
    return GAIL_HIP_OK;
}

uint64_t encoderBytes(GailHipDevice*, const GailHipLayout* L, const GailHipEncodeSettings*) {
    return (uint64_t)L->width * L->height * 16;
}

// One frame's stand-in Tier-1 result: a coding pass in every seventh block, its bytes from a
// hash of the input.
struct FrameResult {
    std::vector<GailHipCblkInfo> info;
    std::vector<uint32_t> passLen;
    std::vector<int32_t> passDist;
    std::vector<uint8_t> data;
    void code(int numCblks, uint64_t h) {
        info.assign(numCblks, GailHipCblkInfo{0, 0, 0, 0, 0});
        passLen.clear();
        passDist.clear();
        data.clear();
        for (int i = 0; i < numCblks; i += 7) {
            GailHipCblkInfo& b = info[i];
            b.nbps = 1;
            b.numPasses = 1;
            b.dataOffset = (uint32_t)data.size();
            b.dataLen = 3;
            b.passBase = (uint32_t)passLen.size();
            for (int k = 0; k < 3; ++k) data.push_back((uint8_t)((h >> (8 * ((i + k) % 8))) & 0x7F));
            passLen.push_back(3);
            passDist.push_back(100 + (int32_t)((h >> (i % 48)) & 63));
        }
    }
    void view(GailHipTier1* out) const {
        out->info = info.data();
        out->passLen = passLen.data();
        out->passDist = passDist.data();
        out->data = data.data();
        out->numInfo = info.size();
        out->numPasses = passLen.size();
        out->numData = data.size();
    }
};

// Bytes of one frame in the input layout the settings describe.
size_t inputBytes(const GailHipEncodeSettings& s) {
    const GailHipFrameFormat& f = s.input;
    if (f.kind != 1) return (size_t)f.width * f.height * 6;
    const size_t cw = (size_t)((f.activeW + f.sx) >> f.sx), ch = (size_t)((f.activeH + f.sy) >> f.sy);
    return ((size_t)f.activeW * f.activeH + 2 * cw * ch) * (f.bits > 8 ? 2 : 1);
}

int encoderCreate(GailHipDevice* dh, const GailHipLayout* L, const GailHipEncodeSettings* s, GailHipEncoder** out) {
    Device* d = reinterpret_cast<Device*>(dh);
    const long limit = envLong("GAIL_STUB_SLOTS", -1);
    if (limit >= 0 && d->encoders.load() >= limit) return failWith(GAIL_HIP_DEVICE_MEMORY, "out of device memory (stand-in)");
    Encoder* e = new Encoder();
    e->dev = d;
    e->numCblks = L->numCblks;
    e->inBytes = inputBytes(*s);
    ++d->encoders;
    *out = reinterpret_cast<GailHipEncoder*>(e);
    return GAIL_HIP_OK;
}

void encoderDestroy(GailHipEncoder* eh) {
    Encoder* e = reinterpret_cast<Encoder*>(eh);
    --e->dev->encoders;
    delete e;
}

void* encoderInput(GailHipEncoder*) { return nullptr; }
void encoderSetSlope(GailHipEncoder* e, double t) { reinterpret_cast<Encoder*>(e)->slope = t; }

// FNV-1a over the first 64 KiB of the input and a sample of words spread over the rest.
uint64_t inputHash(const void* frame, size_t bytes) {
    const uint8_t* p = static_cast<const uint8_t*>(frame);
    uint64_t h = 1469598103934665603ull;
    const size_t head = bytes < 65536 ? bytes : 65536;
    for (size_t i = 0; i < head; ++i) h = (h ^ p[i]) * 1099511628211ull;
    for (size_t i = head; i + 8 <= bytes; i += 4099) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * 1099511628211ull;
    }
    return h;
}

int encoderEncode(GailHipEncoder* eh, const void* frame, GailHipTier1* out, GailHipFrameTimings* timings, GailHipDebug*) {
    Encoder* e = reinterpret_cast<Encoder*>(eh);
    const long n = ++e->dev->encodes;
    const long failAt = envLong("GAIL_STUB_FAIL_AT", 0);
    if (failAt > 0 && n >= failAt) return failWith(GAIL_HIP_ERROR, "device lost (stand-in)");
    const long delay = envLong("GAIL_STUB_DELAY_MS", 0);
    if (delay > 0) std::this_thread::sleep_for(std::chrono::milliseconds(delay));
    if (!frame || !out) return failWith(GAIL_HIP_MISUSE, "null argument");
    const uint64_t h = inputHash(frame, e->inBytes);
    e->info.assign(e->numCblks, GailHipCblkInfo{0, 0, 0, 0, 0});
    e->passLen.clear();
    e->passDist.clear();
    e->data.clear();
    for (int i = 0; i < e->numCblks; i += 7) {
        GailHipCblkInfo& b = e->info[i];
        b.nbps = 1;
        b.numPasses = 1;
        b.dataOffset = (uint32_t)e->data.size();
        b.dataLen = 3;
        b.passBase = (uint32_t)e->passLen.size();
        for (int k = 0; k < 3; ++k) e->data.push_back((uint8_t)((h >> (8 * ((i + k) % 8))) & 0x7F));
        e->passLen.push_back(3);
        e->passDist.push_back(100 + (int32_t)((h >> (i % 48)) & 63));
    }
    out->info = e->info.data();
    out->passLen = e->passLen.data();
    out->passDist = e->passDist.data();
    out->data = e->data.data();
    out->numInfo = e->info.size();
    out->numPasses = e->passLen.size();
    out->numData = e->data.size();
    if (timings) std::memset(timings, 0, sizeof *timings);
    return GAIL_HIP_OK;
}

// ---- batches of frame slots (the served form's jobs)
struct Batch {
    Device* dev;
    uint64_t bytes;
    std::vector<GailHipSlot> slots;
    std::vector<std::vector<uint8_t>> staging;  // one buffer per slot
    struct Frame {
        int numCblks = 0;
        size_t inBytes = 0;
        const void* src = nullptr;
        double slope = 0;
        bool rate = false;
        FrameResult result;
    };
    std::vector<Frame> frames;
    bool launched = false, failed = false;
};

uint64_t slotsBytes(GailHipDevice*, const GailHipSlot* s, int count, const GailHipBatchOptions*) {
    uint64_t b = 0;
    for (int i = 0; i < count; ++i) b += (uint64_t)s[i].maxWidth * s[i].maxHeight * 40;
    return b;
}

int slotsCreate(GailHipDevice* dh, const GailHipSlot* s, int count, const GailHipBatchOptions* o, GailHipBatch** out) {
    Device* d = reinterpret_cast<Device*>(dh);
    const uint64_t need = slotsBytes(dh, s, count, o);
    if (need > deviceAvailable(d)) return failWith(GAIL_HIP_DEVICE_MEMORY, "out of device memory (stand-in)");
    Batch* b = new Batch();
    b->dev = d;
    b->bytes = need;
    b->slots.assign(s, s + count);
    for (int i = 0; i < count; ++i) b->staging.emplace_back((size_t)s[i].maxWidth * s[i].maxHeight * 6);
    d->slotBytes += need;
    *out = reinterpret_cast<GailHipBatch*>(b);
    return GAIL_HIP_OK;
}

void batchDestroy(GailHipBatch* bh) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    b->dev->slotBytes -= b->bytes;
    delete b;
}

int batchBind(GailHipBatch* bh, const GailHipLayout* l, const GailHipEncodeSettings* st, int count) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (count < 1 || count > (int)b->slots.size()) return failWith(GAIL_HIP_MISUSE, "more frames than slots");
    b->frames.assign(count, Batch::Frame{});
    for (int i = 0; i < count; ++i) {
        if (l[i].width > b->slots[i].maxWidth || l[i].height > b->slots[i].maxHeight)
            return failWith(GAIL_HIP_MISUSE, "a frame larger than its slot");
        b->frames[i].numCblks = l[i].numCblks;
        b->frames[i].inBytes = inputBytes(st[i]);
        b->frames[i].rate = st[i].rcMaxTotal != 0;
    }
    return GAIL_HIP_OK;
}

void* batchInput(GailHipBatch* bh, int i) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (i < 0 || i >= (int)b->staging.size()) {
        failWith(GAIL_HIP_MISUSE, "no such slot");
        return nullptr;
    }
    return b->staging[i].data();
}

int batchUploadFrom(GailHipBatch* bh, int i, const void* src) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (i < 0 || i >= (int)b->frames.size() || !src) return failWith(GAIL_HIP_MISUSE, "no such frame");
    b->frames[i].src = src;
    return GAIL_HIP_OK;
}

void batchSetSlope(GailHipBatch* bh, int i, double t) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (i >= 0 && i < (int)b->frames.size()) b->frames[i].slope = t;
}

int batchLaunch(GailHipBatch* bh, int) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    const long delay = envLong("GAIL_STUB_DELAY_MS", 0);
    if (delay > 0) std::this_thread::sleep_for(std::chrono::milliseconds(delay));
    const long failAt = envLong("GAIL_STUB_FAIL_AT", 0);
    b->failed = false;
    for (size_t i = 0; i < b->frames.size(); ++i) {
        Batch::Frame& f = b->frames[i];
        const long n = ++b->dev->encodes;
        if (failAt > 0 && n >= failAt) b->failed = true;
        if (!f.src) return failWith(GAIL_HIP_MISUSE, "a bound frame without input");
        std::memcpy(b->staging[i].data(), f.src, std::min(f.inBytes, b->staging[i].size()));  // the "upload"
        f.result.code(f.numCblks, inputHash(b->staging[i].data(), f.inBytes));
    }
    b->launched = true;
    return GAIL_HIP_OK;
}

int batchWait(GailHipBatch* bh, GailHipTier1* outs, GailHipFrameTimings* timings) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b->launched) return failWith(GAIL_HIP_MISUSE, "nothing launched");
    b->launched = false;
    if (b->failed) return failWith(GAIL_HIP_ERROR, "device lost (stand-in)");
    for (size_t i = 0; i < b->frames.size(); ++i) b->frames[i].result.view(&outs[i]);
    if (timings) std::memset(timings, 0, sizeof *timings);
    return GAIL_HIP_OK;
}

int batchQuery(GailHipBatch*) { return 1; }
int batchInputsReady(GailHipBatch*) { return GAIL_HIP_OK; }  // stand-in uploads are synchronous

// Device rate control of a stand-in frame: every pass is kept (its few bytes fit any budget).
int batchRate(GailHipBatch* bh, int i, GailHipRate* out) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (i < 0 || i >= (int)b->frames.size() || !b->frames[i].rate)
        return failWith(GAIL_HIP_MISUSE, "the frame had no rate control on the device");
    std::memset(out, 0, sizeof *out);
    out->lambda[0] = out->lambda[1] = out->lambda[2] = 1.0;
    return GAIL_HIP_OK;
}

void* hostAlloc(uint64_t bytes) { return std::malloc(bytes ? bytes : 16); }
void hostFree(void* p) { std::free(p); }
int isDevicePointer(const void*) { return 0; }

int unsupported() { return failWith(GAIL_HIP_UNSUPPORTED, "a stand-in plugin has no decoder"); }
uint64_t decoderBytes(GailHipDevice*, const GailHipLayout*, const GailHipDecodeSettings*) { return 0; }
int decoderCreate(GailHipDevice*, const GailHipLayout*, const GailHipDecodeSettings*, GailHipDecoder**) { return unsupported(); }
void decoderDestroy(GailHipDecoder*) {}
int decoderDecode(GailHipDecoder*, const GailHipLayout*, const uint8_t*, size_t, const GailHipCblkCode*, size_t, void*, ptrdiff_t,
                  GailHipDecodeTimings*) {
    return unsupported();
}

GailHipPlugin makeTable() {
    GailHipPlugin t;
    std::memset(&t, 0, sizeof t);
    t.abi = STUB_MODE == 1 ? GAIL_HIP_PLUGIN_ABI + 1 : GAIL_HIP_PLUGIN_ABI;
    t.size = STUB_MODE == 3 ? (uint32_t)sizeof(GailHipPlugin) - 8 : (uint32_t)sizeof(GailHipPlugin);
    t.layout = STUB_MODE == 4 ? gail_hip_plugin_layout() ^ 1 : gail_hip_plugin_layout();
    t.name = "gail_hip";
    t.toolkit = "0.0 (stand-in)";
    t.archs = "gfx950";
    t.last_error = lastError;
    t.device_count = deviceCount;
    t.device_desc = deviceDesc;
    t.device_usable = deviceUsable;
    t.device_create = deviceCreate;
    t.device_destroy = deviceDestroy;
    t.device_name = deviceName;
    t.device_memory = deviceMemory;
    t.encoder_bytes = encoderBytes;
    t.encoder_create = encoderCreate;
    t.encoder_destroy = encoderDestroy;
    t.encoder_input = encoderInput;
    t.encoder_set_slope = encoderSetSlope;
    t.encoder_encode = encoderEncode;
    t.decoder_bytes = decoderBytes;
    t.decoder_create = decoderCreate;
    t.decoder_destroy = decoderDestroy;
    t.decoder_decode = decoderDecode;
    t.batch_destroy = batchDestroy;
    t.batch_input = batchInput;
    t.batch_set_slope = batchSetSlope;
    t.batch_launch = batchLaunch;
    t.batch_wait = batchWait;
    t.slots_bytes = slotsBytes;
    t.slots_create = slotsCreate;
    t.batch_bind = batchBind;
    t.batch_query = batchQuery;
    t.batch_upload_from = batchUploadFrom;
    t.host_alloc = hostAlloc;
    t.host_free = hostFree;
    t.is_device_pointer = isDevicePointer;
    t.batch_rate = batchRate;
    t.batch_inputs_ready = batchInputsReady;
    return t;  // the entries the served form does not use stay NULL
}

const GailHipPlugin kTable = makeTable();

}  // namespace

// Sets one of the GAIL_STUB_* knobs from the test program (found with dlopen RTLD_NOLOAD on
// the plugin libgail.so loaded); returns 0, or -1 for an unknown name.
extern "C" __attribute__((visibility("default"))) int gail_stub_control(const char* name, long value) {
    for (Knob& k : g_knobs)
        if (!std::strcmp(k.name, name)) {
            k.value.store(value);
            k.set.store(true);
            return 0;
        }
    return -1;
}

#if STUB_MODE != 5
extern "C" __attribute__((visibility("default"))) const GailHipPlugin* gail_hip_plugin_get(uint32_t abi) {
    if (STUB_MODE == 2) return nullptr;
    if (STUB_MODE == 1) return &kTable;
    return abi == GAIL_HIP_PLUGIN_ABI ? &kTable : nullptr;
}
#endif
