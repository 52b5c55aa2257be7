// Host-built tables that make the device colour path and the Tier-1 early-termination
// estimate reproduce the CPU reference's float results exactly.
#pragma once
#include <cstdint>
#include <vector>

namespace gail_amd {

// Colour handling of R'G'B' / Y'CbCr sources (SDR; HDR sources in the default "signal"
// mode, which uses the gamma-2.4 table). transfer: 0 BT.1886, 1 sRGB, 2 PQ, 3 HLG,
// 4 linear, 5 gamma 2.6; primaries: 0 BT.709, 1 BT.2020, 2 P3-D65, 3 DCI-P3, 4 X'Y'Z'.
struct ColorSpec {
    int transfer = 0, primaries = 0;
    double sdrWhite = 48.0, cinemaPeak = 48.0;
};

// The 65,536-entry EOTF table in cd/m^2 (color.cpp:296-326) and the 3x4 linear RGB ->
// XYZ matrix rows (color.cpp:59-79, 327-330).
void buildColorTables(const ColorSpec& c, std::vector<float>& lut, float matrix[12]);

// thresholds[k] (k = 1..4095): the smallest float v with
// floor(4095 * powf(clamp(v / 52.37, 0, 1), 1 / 2.6) + 0.5) >= k, using this host's powf
// (color_math.h:109-112). That function of v is monotone, so counting thresholds <= v
// gives its exact value. thresholds[0] = 0.
void buildCodeThresholds(std::vector<float>& thresholds);

// log2f(max(k, 0.5) / n) for 1 <= n <= maxN, 0 <= k <= n at index n * (n + 1) / 2 + k,
// with this host's log2f: every argument the per-plane entropy estimate of the early
// termination can see (t1_ref.cpp:128-140).
void buildLog2Table(std::vector<float>& table, uint32_t maxN);

}  // namespace gail_amd
