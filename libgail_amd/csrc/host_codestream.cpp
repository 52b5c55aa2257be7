// C entry points of libgail.so for the host half of a frame's encode: PCRD-opt rate
// allocation and Tier-2 assembly of one frame's Tier-1 output into a DCI codestream
// (what the reference's pipeline worker does after EncoderSlot::encode,
// pipeline.cpp:210-226), and the DCI structural check.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <algorithm>
#include <string>
#include <vector>

#include "codestream.h"
#include "gail_amd_export.h"
#include "layout.h"

using namespace gail_amd;

namespace {

const Layout& cachedLayout(int W, int H, bool fourk, double step) {
    static std::mutex m;
    static std::map<std::tuple<int, int, bool, double>, std::unique_ptr<Layout>> cache;
    std::lock_guard<std::mutex> l(m);
    auto& p = cache[std::make_tuple(W, H, fourk, step)];
    if (!p) p = std::make_unique<Layout>(makeLayout(W, H, fourk, step));
    return *p;
}

}  // namespace

// Rate control (budget from fps and Mbit/s, 200 Mbit/s per component) and, when out is
// given, the codestream itself. Returns the codestream size (the measured size when out
// is NULL), or -1 on an error / -2 when out is too small. lambda receives the smallest of
// the per-component thresholds (the value the pipeline feeds to the next frame's early
// termination), mse the estimated squared error per sample in 12-bit code units.
extern "C" GAIL_AMD_EXPORT int64_t gail_encode_codestream(int W, int H, int fourk, double baseStep, const uint32_t* info5,
                                                          const uint32_t* passLen, const int32_t* passDist,
                                                          const uint8_t* data, uint64_t nblocks, double fps,
                                                          double mbps, uint8_t* out, uint64_t cap, double* lambda,
                                                          double* mse, const char* tag) {
    try {
        const Layout& L = cachedLayout(W, H, fourk != 0, baseStep);
        if ((uint64_t)L.numCblks() != nblocks) return -1;
        const Tier1Ref t1{info5, passLen, passDist, data, (size_t)nblocks};
        const RateResult rr = allocateRate(L, t1, dciLimits(fps, mbps, 200.0));
        if (lambda) *lambda = std::min({rr.lambdaComp[0], rr.lambdaComp[1], rr.lambdaComp[2]});
        if (mse) *mse = rr.distortion / ((double)W * H * 3);
        if (!out) return (int64_t)rr.sizes.total;
        std::vector<uint8_t> cs;
        buildCodestream(L, t1, rr.passes.data(), &cs, nullptr, tag ? tag : "");
        if (cs.size() > cap) return -2;
        std::memcpy(out, cs.data(), cs.size());
        return (int64_t)cs.size();
    } catch (...) {
        return -1;
    }
}

// The rate-control slope grid (floor(1024 log2 s) and a bin's lowest double), for tests.
extern "C" GAIL_AMD_EXPORT int64_t gail_slope_bin(double s) { return s > 0 && std::isfinite(s) ? slopeBin(s) : 0; }
extern "C" GAIL_AMD_EXPORT double gail_slope_bin_edge(int64_t k) { return binEdge(k); }

// "" (returned as 0) when the codestream passes, else a message copied into msg.
extern "C" GAIL_AMD_EXPORT int gail_check_codestream(const uint8_t* data, uint64_t size, double fps, int fourk,
                                                     char* msg, int cap) {
    const std::string r = checkDciCodestream(data, size, fps, fourk != 0);
    if (msg && cap > 0) {
        std::strncpy(msg, r.c_str(), cap - 1);
        msg[cap - 1] = 0;
    }
    return r.empty() ? 0 : 1;
}

extern "C" GAIL_AMD_EXPORT uint64_t gail_empty_codestream_bytes(int W, int H, int fourk, double baseStep) {
    try {
        return emptyCodestreamBytes(cachedLayout(W, H, fourk != 0, baseStep));
    } catch (...) {
        return 0;
    }
}
