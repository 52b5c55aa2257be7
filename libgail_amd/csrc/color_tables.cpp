#include "color_tables.h"

#include <cmath>
#include <cstring>

namespace gail_amd {

namespace {

void invert3(const double m[9], double r[9]) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    r[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    r[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    r[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    r[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    r[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    r[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    r[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    r[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    r[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

// Linear RGB (white Y = 1) -> XYZ from chromaticities: P's columns scaled so that
// RGB = (1, 1, 1) lands on the white point.
void rgbToXyz(int primaries, double M[9]) {
    static const double xy[4][4][2] = {
        {{0.640, 0.330}, {0.300, 0.600}, {0.150, 0.060}, {0.3127, 0.3290}},
        {{0.708, 0.292}, {0.170, 0.797}, {0.131, 0.046}, {0.3127, 0.3290}},
        {{0.680, 0.320}, {0.265, 0.690}, {0.150, 0.060}, {0.3127, 0.3290}},
        {{0.680, 0.320}, {0.265, 0.690}, {0.150, 0.060}, {0.314, 0.351}},
    };
    if (primaries < 0 || primaries > 3) {
        for (int i = 0; i < 9; ++i) M[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    const auto& c = xy[primaries];
    double P[9], Pi[9], S[3];
    for (int i = 0; i < 3; ++i) {
        P[i] = c[i][0] / c[i][1];
        P[3 + i] = 1.0;
        P[6 + i] = (1.0 - c[i][0] - c[i][1]) / c[i][1];
    }
    const double W[3] = {c[3][0] / c[3][1], 1.0, (1.0 - c[3][0] - c[3][1]) / c[3][1]};
    invert3(P, Pi);
    for (int i = 0; i < 3; ++i) S[i] = Pi[3 * i] * W[0] + Pi[3 * i + 1] * W[1] + Pi[3 * i + 2] * W[2];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) M[3 * r + k] = P[3 * r + k] * S[k];
}

float encodeCode(float v) {
    const float n = std::fmin(std::fmax(v / 52.37f, 0.0f), 1.0f);
    return std::floor(4095.0f * std::pow(n, 1.0f / 2.6f) + 0.5f);
}

}  // namespace

void buildColorTables(const ColorSpec& c, std::vector<float>& lut, float matrix[12]) {
    lut.resize(65536);
    for (int i = 0; i < 65536; ++i) {
        const double e = i / 65535.0;
        double nits;
        switch (c.transfer) {
        case 1: nits = (e <= 0.04045 ? e / 12.92 : std::pow((e + 0.055) / 1.055, 2.4)) * c.sdrWhite; break;
        case 4: nits = e * c.sdrWhite; break;
        case 5: nits = std::pow(e, 2.6) * 52.37; break;
        default: nits = std::pow(e, 2.4) * c.sdrWhite; break;
        }
        lut[i] = (float)nits;
    }
    double M[9];
    rgbToXyz(c.primaries, M);
    std::memset(matrix, 0, 12 * sizeof(float));
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) matrix[4 * r + k] = (float)M[3 * r + k];
}

void buildCodeThresholds(std::vector<float>& thr) {
    thr.assign(4096, 0.f);
    uint32_t hiBits;
    const float top = 64.0f;
    std::memcpy(&hiBits, &top, 4);
    for (int k = 1; k < 4096; ++k) {
        // smallest non-negative float (ordered like its bit pattern) with code >= k
        uint32_t lo = 0, hi = hiBits;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float v;
            std::memcpy(&v, &mid, 4);
            if (encodeCode(v) >= (float)k) hi = mid;
            else lo = mid + 1;
        }
        std::memcpy(&thr[k], &lo, 4);
    }
}

void buildLog2Table(std::vector<float>& t, uint32_t maxN) {
    t.assign((size_t)(maxN + 1) * (maxN + 2) / 2, 0.f);
    for (uint32_t n = 1; n <= maxN; ++n) {
        const float fn = (float)n;
        float* row = &t[(size_t)n * (n + 1) / 2];
        for (uint32_t k = 0; k <= n; ++k) {
            const float fk = (float)k;
            row[k] = std::log2(std::max(fk, 0.5f) / fn);
        }
    }
}

}  // namespace gail_amd
