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
//   GAIL_STUB_FAIL_AT=n    the n-th encode on a device (1-based) and every later one fail as
//                          a lost device would (GAIL_HIP_ERROR)
//   GAIL_STUB_SLOTS=k      encoder_create reports device memory exhausted after k encoders
//   GAIL_STUB_DELAY_MS=d   every encode takes d ms longer
#include <stdlib.h>

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
    const char* v = std::getenv(name);
    return v && *v ? std::strtol(v, nullptr, 10) : dflt;
}

int failWith(int status, const char* msg) {
    g_error = msg;
    return status;
}

struct Device {
    std::string name = "Stand-in MI355X (HIP test plugin)";
    std::atomic<long> encodes{0};
    std::atomic<int> encoders{0};
};

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

int deviceMemory(GailHipDevice*, GailHipMemory* out) {
    std::memset(out, 0, sizeof *out);  // not budgeted: the library takes frames_in_flight as asked
    return GAIL_HIP_OK;
}

uint64_t encoderBytes(GailHipDevice*, const GailHipLayout* L, const GailHipEncodeSettings*) {
    return (uint64_t)L->width * L->height * 16;
}

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

int unsupported() { return failWith(GAIL_HIP_UNSUPPORTED, "a stand-in plugin has no decoder or batches"); }
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
    return t;  // batch entries stay NULL: libgail's encoder does not use them
}

const GailHipPlugin kTable = makeTable();

}  // namespace

#if STUB_MODE != 5
extern "C" __attribute__((visibility("default"))) const GailHipPlugin* gail_hip_plugin_get(uint32_t abi) {
    if (STUB_MODE == 2) return nullptr;
    if (STUB_MODE == 1) return &kTable;
    return abi == GAIL_HIP_PLUGIN_ABI ? &kTable : nullptr;
}
#endif
