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
        buildCodestream(L, t1, rr.passes.data(), &cs, nullptr, tag ? tag : "", true);
        if (cs.size() > cap) return -2;
        std::memcpy(out, cs.data(), cs.size());
        return (int64_t)cs.size();
    } catch (...) {
        return -1;
    }
}

// Sizes of the same frame measured two ways, buildCodestream and the device's packet-size
// code run on the host, over `trials` pass choices (all passes, none, and random truncations
// from seed); returns the number of choices where any size differs, or -1 on error.
extern "C" GAIL_AMD_EXPORT int gail_rc_measure_check(int W, int H, int fourk, double baseStep, const uint32_t* info5,
                                                     const uint32_t* passLen, const int32_t* passDist,
                                                     const uint8_t* data, uint64_t nblocks, int trials, uint64_t seed) {
    try {
        const Layout& L = cachedLayout(W, H, fourk != 0, baseStep);
        if ((uint64_t)L.numCblks() != nblocks) return -1;
        const Tier1Ref t1{info5, passLen, passDist, data, (size_t)nblocks};
        const RcTables T = makeRcTables(L);
        std::vector<uint8_t> passes(nblocks);
        uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
        int bad = 0;
        for (int t = 0; t < trials; ++t) {
            for (uint64_t i = 0; i < nblocks; ++i) {
                const uint32_t np = t1.passes((int)i);
                x ^= x << 13;
                x ^= x >> 7;
                x ^= x << 17;
                // all passes, none, random truncations, and sparse inclusion (about one block
                // in four) that leaves whole quads and precinct-bands empty
                const uint32_t r = np ? (uint32_t)(x % (np + 1)) : 0u;
                passes[i] = (uint8_t)(t == 0 ? np : t == 1 ? 0 : (t & 1) ? r : ((x >> 40) % 4 == 0 ? r : 0u));
            }
            CodestreamSizes a, b = rcMeasure(L, T, t1, passes.data());
            buildCodestream(L, t1, passes.data(), nullptr, &a);
            bool same = a.total == b.total && a.numTileParts == b.numTileParts;
            for (int c = 0; c < 3; ++c) same = same && a.comp[c] == b.comp[c];
            for (int k = 0; k < a.numTileParts; ++k) same = same && a.tilePart[k] == b.tilePart[k];
            if (!same) ++bad;
        }
        return bad;
    } catch (...) {
        return -1;
    }
}

// Rate control of one frame against explicit byte budgets (maxTotal for the whole codestream,
// maxComp per component): passes[nblocks] receives the passes kept per block, lambda[3] the
// per-component thresholds, sizes[4] the codestream's total and component bytes as Tier-2 will
// write them, distortion the weighted squared error left. Returns the iterations, or -1.
extern "C" GAIL_AMD_EXPORT int gail_rate_allocate(int W, int H, int fourk, double baseStep, const uint32_t* info5,
                                                  const uint32_t* passLen, const int32_t* passDist, const uint8_t* data,
                                                  uint64_t nblocks, uint64_t maxTotal, uint64_t maxComp, uint8_t* passes,
                                                  double* lambda, uint64_t* sizes, double* distortion) {
    try {
        const Layout& L = cachedLayout(W, H, fourk != 0, baseStep);
        if ((uint64_t)L.numCblks() != nblocks) return -1;
        const Tier1Ref t1{info5, passLen, passDist, data, (size_t)nblocks};
        const RateResult rr = allocateRate(L, t1, RateLimits{(size_t)maxTotal, (size_t)maxComp});
        std::memcpy(passes, rr.passes.data(), nblocks);
        for (int c = 0; c < 3; ++c) lambda[c] = rr.lambdaComp[c];
        sizes[0] = rr.sizes.total;
        for (int c = 0; c < 3; ++c) sizes[1 + c] = rr.sizes.comp[c];
        *distortion = rr.distortion;
        return rr.iterations;
    } catch (...) {
        return -1;
    }
}

// dciLimits(fps, maxMbps, compMbps) as two byte budgets.
extern "C" GAIL_AMD_EXPORT void gail_dci_limits(double fps, double maxMbps, double compMbps, uint64_t* total,
                                                uint64_t* comp) {
    const RateLimits l = dciLimits(fps, maxMbps, compMbps);
    *total = l.maxTotal;
    *comp = l.maxComp;
}

// The rate-control slope grid (floor(1024 log2 s) and a bin's lowest double), for tests.
extern "C" GAIL_AMD_EXPORT int64_t gail_slope_bin(double s) { return s > 0 && std::isfinite(s) ? slopeBin(s) : 0; }
extern "C" GAIL_AMD_EXPORT double gail_slope_bin_edge(int64_t k) { return binEdge(k); }

// Tier-2 of a frame whose passes were already chosen (each block keeps t1.passes(i), as the
// device rate control leaves them): the codestream bytes, or -1 / -2 (too small a buffer).
extern "C" GAIL_AMD_EXPORT int64_t gail_build_codestream(int W, int H, int fourk, double baseStep, const uint32_t* info5,
                                                         const uint32_t* passLen, const int32_t* passDist,
                                                         const uint8_t* data, uint64_t nblocks, uint8_t* out, uint64_t cap,
                                                         const char* tag) {
    try {
        const Layout& L = cachedLayout(W, H, fourk != 0, baseStep);
        if ((uint64_t)L.numCblks() != nblocks) return -1;
        const Tier1Ref t1{info5, passLen, passDist, data, (size_t)nblocks};
        std::vector<uint8_t> passes(nblocks);
        for (uint64_t i = 0; i < nblocks; ++i) passes[i] = (uint8_t)t1.passes((int)i);
        std::vector<uint8_t> cs;
        buildCodestream(L, t1, passes.data(), &cs, nullptr, tag ? tag : "", true);
        if (cs.size() > cap) return -2;
        std::memcpy(out, cs.data(), cs.size());
        return (int64_t)cs.size();
    } catch (...) {
        return -1;
    }
}

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
