// TEST INFRASTRUCTURE. The reference library's device interfaces driven through the HIP
// adapter (libgail_amd/adapter/hip_host.cpp), built with the reference's own sources
// (oracle/Makefile, `make ref`) into oracle/_ref/adapter_check:
//
//   adapter_check <plugin dir> load      the devices createHipDevices sets up
//   adapter_check <plugin dir> encode    a 2K flat X'Y'Z' frame (RGB48 as the reference's API
//                                        hands it over) through HipEncoderSlot and through the
//                                        reference's own CPU slot (cpu_backend.cpp): Tier-1
//                                        output equal block for block, at two early-
//                                        termination thresholds
//   adapter_check <plugin dir> decode    that frame's Tier-1 output decoded by HipDecoderSlot
//                                        and by the reference's CPU decode slot: equal pictures
//   adapter_check <plugin dir> shape     slots made and a frame coded (the Tier-1 records
//                                        cover the layout), no comparison: for stand-ins
//   adapter_check <plugin dir> nomem     createSlot throws gail::DeviceOutOfMemory
//
// Prints what it checked; exit status 0 when everything held.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "core/backend.h"
#include "core/thread_pool.h"
#include "hip_host.h"

using namespace gail;

namespace {

int failures = 0;
#define EXPECT(c, ...)                               \
    do {                                             \
        if (!(c)) {                                  \
            std::printf("  FAIL line %d: ", __LINE__); \
            std::printf(__VA_ARGS__);                \
            std::printf("\n");                       \
            ++failures;                              \
        }                                            \
    } while (0)

// 12-bit X'Y'Z' codes of a deterministic picture, widened to 16 bits the way the reference's
// gail_api.cpp:389-394 hands them to a backend, interleaved over the whole container.
std::vector<uint16_t> xyzFrame(int W, int H, uint32_t seed) {
    std::vector<uint16_t> px((size_t)W * H * 3);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c) {
                uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)c * 83492791u ^ seed;
                h ^= h >> 13;
                h *= 0x5bd1e995u;
                const uint32_t noise = (h ^ (h >> 15)) % 61u;
                const uint32_t v = (uint32_t)((x * 3 + y * 2 + c * 700 + (x / 37) * (y / 29) * 11) % 3800) + noise;
                const uint32_t code = v > 4095 ? 4095 : v;
                px[((size_t)y * W + x) * 3 + c] = (uint16_t)((code * 65535u + 2047u) / 4095u);
            }
    return px;
}

EncodeSettings xyzSettings(const FrameLayout& L) {
    EncodeSettings s;
    s.profile = L.profile;
    s.width = L.width;
    s.height = L.height;
    s.baseStep = L.baseStep;
    s.input.kind = FrameFormat::Kind::RGB48;
    s.input.width = L.width;
    s.input.height = L.height;
    s.color.transfer = Transfer::Gamma26;  // X'Y'Z' passes through untouched
    s.color.primaries = Primaries::XYZ;
    return s;
}

int compareTier1(const Tier1Frame& a, const Tier1Frame& b) {
    if (a.info.size() != b.info.size()) return (int)std::max(a.info.size(), b.info.size());
    int bad = 0;
    for (size_t i = 0; i < a.info.size(); ++i) {
        const CblkInfo &x = a.info[i], &y = b.info[i];
        bool same = x.nbps == y.nbps && x.numPasses == y.numPasses && x.dataLen == y.dataLen;
        for (uint32_t k = 0; same && k < x.numPasses; ++k)
            same = a.len((int)i, (int)k) == b.len((int)i, (int)k) && a.dist((int)i, (int)k) == b.dist((int)i, (int)k);
        same = same && std::memcmp(a.data.data() + x.dataOffset, b.data.data() + y.dataOffset, x.dataLen) == 0;
        bad += same ? 0 : 1;
    }
    return bad;
}

std::unique_ptr<Device> hipDevice(const std::string& dir) {
    std::vector<std::unique_ptr<Device>> devs = createHipDevices(dir, {}, [](LogLevel, const std::string& m) {
        std::printf("  log: %s\n", m.c_str());
    });
    return devs.empty() ? nullptr : std::move(devs[0]);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: adapter_check <plugin dir> load|encode|decode|shape|nomem\n");
        return 2;
    }
    const std::string dir = argv[1], mode = argv[2];
    const HipPluginInfo info = loadHipPlugin(dir);
    if (mode == "load") {
        std::unique_ptr<Device> d = hipDevice(dir);
        std::printf("  plugin %s: %s\n", info.path.c_str(), info.problem.empty() ? "loaded" : info.problem.c_str());
        if (d) std::printf("  device %s, tag \"%s\"\n", d->name().c_str(), d->codestreamTag().c_str());
        EXPECT(d != nullptr, "no device");
        return failures ? 1 : 0;
    }
    std::unique_ptr<Device> hip = hipDevice(dir);
    if (!hip) {
        std::printf("  no HIP device: %s\n", info.problem.c_str());
        return 1;
    }
    const FrameLayout L = makeLayout(1998, 1080, Profile::Cinema2K, 1.0);
    const EncodeSettings s = xyzSettings(L);
    const std::vector<uint16_t> frame = xyzFrame(L.width, L.height, 12345u);
    try {
        if (mode == "nomem") {
            bool threw = false;
            try {
                hip->createSlot(L, s);
            } catch (const DeviceOutOfMemory& e) {
                threw = true;
                std::printf("  createSlot: DeviceOutOfMemory (%s)\n", e.what());
            }
            EXPECT(threw, "createSlot did not throw DeviceOutOfMemory");
        } else if (mode == "shape") {
            std::unique_ptr<EncoderSlot> slot = hip->createSlot(L, s);
            Tier1Frame t;
            slot->encode(frame.data(), t);
            EXPECT((int)t.info.size() == L.numCblks(), "%zu records for %d code-blocks", t.info.size(), L.numCblks());
            std::printf("  slot made, frame coded: %zu records, %zu bytes; %.0f MB per slot, %s memory\n", t.info.size(),
                        t.data.size(), hip->slotBytes(L, s) / 1e6, hip->memory().budgeted ? "budgeted" : "unknown");
        } else if (mode == "encode" || mode == "decode") {
            std::unique_ptr<Device> cpu = createCpuDevice(8);
            std::unique_ptr<EncoderSlot> a = hip->createSlot(L, s), b = cpu->createSlot(L, s);
            Tier1Frame ta, tb;
            for (double thr : {0.0, 60.0}) {
                a->setSlopeThreshold(thr);
                b->setSlopeThreshold(thr);
                FrameTimings tm;
                a->encode(frame.data(), ta, &tm);
                b->encode(frame.data(), tb);
                const int bad = compareTier1(ta, tb);
                std::printf("  threshold %g: %zu code-blocks, %zu / %zu codeword bytes (HIP / reference CPU slot), "
                            "%d differ; HIP Tier-1 %.2f ms\n",
                            thr, ta.info.size(), ta.data.size(), tb.data.size(), bad, tm.t1);
                EXPECT(bad == 0, "%d code-blocks differ from the reference's CPU slot", bad);
            }
            if (mode == "decode") {
                // the Tier-1 output read as its own codestream: every block's bytes at its offset
                std::vector<CblkCode> blocks(ta.info.size());
                for (size_t i = 0; i < blocks.size(); ++i) {
                    blocks[i].offset = ta.info[i].dataOffset;
                    blocks[i].len = ta.info[i].dataLen;
                    blocks[i].passes = (uint8_t)std::min<uint32_t>(ta.info[i].numPasses, 255);
                    blocks[i].nbps = (uint8_t)ta.info[i].nbps;
                }
                for (int reduce : {0, 2}) {
                    DecodeSettings ds;
                    ds.reduce = reduce;
                    std::unique_ptr<DecoderSlot> da = hip->createDecoderSlot(L, ds);
                    std::unique_ptr<DecoderSlot> db = createCpuDecoderSlot(L, ds, std::make_shared<ThreadPool>(8, false));
                    int w, h;
                    decodedSize(L, reduce, w, h);
                    std::vector<uint16_t> pa((size_t)w * h * 3, 1), pb((size_t)w * h * 3, 2);
                    da->decode(L, ta.data.data(), ta.data.size(), blocks, pa.data(), (ptrdiff_t)w * 6, nullptr);
                    db->decode(L, ta.data.data(), ta.data.size(), blocks, pb.data(), (ptrdiff_t)w * 6, nullptr);
                    size_t diff = 0;
                    for (size_t i = 0; i < pa.size(); ++i) diff += pa[i] != pb[i];
                    std::printf("  decode reduce %d: %dx%d picture, %zu words differ from the reference's CPU slot\n",
                                reduce, w, h, diff);
                    EXPECT(diff == 0, "decoded pictures differ");
                }
            }
        } else {
            std::fprintf(stderr, "adapter_check: unknown mode %s\n", mode.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("  FAIL: exception %s\n", e.what());
        return 1;
    }
    std::printf("%s: %d failure(s)\n", mode.c_str(), failures);
    return failures ? 1 : 0;
}
