// Tier-2 and rate control on the host: DCI codestream assembly (packet headers with tag
// trees, CPRL progression with the 4K POC, one tile-part per component and progression
// volume, TLM) and PCRD-opt rate allocation under the DCI byte limits.
// Follows the reference's t2.cpp:31-254, progression.h:28-46, tagtree.h, bitio.h and
// rate_control.cpp:11-279 (ISO/IEC 15444-1 Annex B, J.14).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "device_types.h"
#include "layout.h"

namespace gail_amd {

// One frame's Tier-1 output as the backend leaves it (gail::Tier1Frame in plain form).
struct Tier1Ref {
    const uint32_t* info5;    // per block: nbps, passes, data offset, data length, pass base
    const uint32_t* passLen;  // truncation length after each pass
    const int32_t* passDist;  // distortion reduction of each pass (x4096, bit-plane units)
    const uint8_t* data;
    size_t numBlocks;

    uint32_t nbps(int i) const { return info5[5 * i]; }
    uint32_t passes(int i) const { return info5[5 * i + 1]; }
    const uint8_t* code(int i) const { return data + info5[5 * i + 2]; }
    uint32_t len(int i, int k) const { return passLen[info5[5 * i + 4] + k]; }
    int32_t dist(int i, int k) const { return passDist[info5[5 * i + 4] + k]; }
};

struct CodestreamSizes {
    size_t total = 0;
    size_t comp[3] = {0, 0, 0};  // bytes per component (its tile-parts with SOT and SOD)
    int numTileParts = 0;
    size_t tilePart[6] = {};
};

// Longest device tag a codestream's COM marker can carry after the label.
constexpr size_t kMaxDeviceTagBytes = 9;
const char* codestreamLabel();

// Writes (out != nullptr) or only measures the codestream with passes[i] coding passes of
// block i. The sizes count the COM marker at its longest, so the coded data never depends
// on the tag. parallel: packets are coded on the rate-control helper threads as well (one
// frame at a time callers); the bytes are the same either way.
void buildCodestream(const Layout& L, const Tier1Ref& t1, const uint8_t* passes, std::vector<uint8_t>* out,
                     CodestreamSizes* sizes, const std::string& tag = std::string(), bool parallel = false);
// Smallest codestream this layout can produce (no pass anywhere).
size_t emptyCodestreamBytes(const Layout& L);

struct RateLimits {
    size_t maxTotal = 0;  // bytes per frame
    size_t maxComp = 0;   // bytes per component
};
// Budgets for a frame rate (rate_control.cpp:11-26): floor(Mbit/s * 1e6 / (8 fps)) - 16.
RateLimits dciLimits(double fps, double maxMbps = 230.0, double compMbps = 200.0);

struct RateResult {
    std::vector<uint8_t> passes;  // passes included per code-block
    double lambda = 0;
    double lambdaComp[3] = {0, 0, 0};
    double distortion = 0;        // estimated weighted squared error left
    CodestreamSizes sizes;
    int iterations = 0;
};

// The 1,024 steps of one octave as the 52-bit mantissas of exp2(j / 1024), j = 0..1023
// (entry 1024 = 2^52): the grid rate control bins slopes on.
struct OctaveSteps {
    unsigned long long mantissa[1025];
    uint16_t first[1024];  // largest j with mantissa[j] <= r << 42, for r = 0..1023
};
const OctaveSteps& octaveSteps();
// floor(1024 log2 s) for a finite s > 0, and the smallest double of that bin.
int64_t slopeBin(double s);
double binEdge(int64_t k);

// A layout's packets for the device rate control: every packet with its precinct-bands
// and tile-part, and the bytes that do not depend on the coded data.
struct RcTables {
    std::vector<RcPacket> packets;
    std::vector<RcPBand> pbands;
    uint32_t ntp = 0;
    int tpComp[6] = {};
    unsigned long long fixedBytes = 0;  // main header, COM reserve, EOC
};
RcTables makeRcTables(const Layout& L);
// Codestream sizes from the tables and rc_packet_bytes (what the device computes) for
// passes[i] passes of block i.
CodestreamSizes rcMeasure(const Layout& L, const RcTables& T, const Tier1Ref& t1, const uint8_t* passes);

// PCRD-opt under the limits (rate_control.cpp:165-279).
RateResult allocateRate(const Layout& L, const Tier1Ref& t1, const RateLimits& lim);

// Distortion scale of pass k of a block with nbps bit-planes (t1_ref.h:44-47).
inline double passPlaneScale(int nbps, int k) {
    const int plane = k == 0 ? nbps - 1 : nbps - 2 - (k - 1) / 3;
    return (double)(1ull << (2 * plane)) / 4096.0;
}

// Structural check of a DCI codestream (codestream_check.cpp:43-280, the parts a frame of
// this encoder can get wrong): marker order, SIZ / COD / QCD fields of the profile, tile-part
// lengths against TLM and Psot, and the byte limits. Returns "" when it passes.
std::string checkDciCodestream(const uint8_t* data, size_t size, double fps, bool fourk);

}  // namespace gail_amd
