// Host-built tables that make the device colour path and the Tier-1 early-termination
// estimate reproduce the CPU reference's float results exactly.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace gail_amd {

// Colour handling of R'G'B' / Y'CbCr sources (color.h:84-100, in ColorParams'
// enumerations). transfer: 0 BT.1886, 1 sRGB, 2 PQ, 3 HLG, 4 linear, 5 gamma 2.6;
// primaries: 0 BT.709, 1 BT.2020, 2 P3-D65, 3 DCI-P3, 4 X'Y'Z'; toneMap (PQ / HLG sources):
// 0 signal (gamma 2.4 at sdrWhite), 1 shoulder, 2 BT.2390, 3 none; toneNorm: 0 each channel,
// 1 max(R,G,B), 2 luminance; gamut: 0 P3-D65, 1 DCI-P3, 2 none (no compression). An app 3D
// LUT (lut3dSize nodes per axis, red fastest, R'G'B' -> X'Y'Z' in 0..1) replaces all of it.
struct ColorSpec {
    int transfer = 0, primaries = 0, toneMap = 0, toneNorm = 0, gamut = 2;
    double sdrWhite = 48.0, hdrPeak = 1000.0, cinemaPeak = 48.0, hdrRefWhite = 203.0, cinemaRefWhite = 16.0;
    double knee = 0.25, gamutThreshold = 0.8;
    int lut3dSize = 0;
    std::shared_ptr<const std::vector<float>> lut3d;
    // true when both give the same tables (the 3D LUT compared by content)
    bool sameTables(const ColorSpec& o) const;
};

constexpr int kColorParams = 48;  // parameter block of the per-pixel colour path

// The 65,536-entry EOTF table in cd/m^2 and the parameter block (color.cpp:296-372,
// color_params.h:7-23): linear RGB -> XYZ matrix, tone map, tone norm, target-gamut
// matrices and compression, 3D LUT size.
void buildColorTables(const ColorSpec& c, std::vector<float>& lut, float params[kColorParams]);

// An app 3D LUT from a .cube file (LUT_3D_SIZE, domain 0..1; color.cpp:240-279).
std::vector<float> readCubeLut(const std::string& path, int& size);

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
