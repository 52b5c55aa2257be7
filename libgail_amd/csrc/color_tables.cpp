#include "color_tables.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>

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

bool ColorSpec::sameTables(const ColorSpec& o) const {
    if (transfer != o.transfer || primaries != o.primaries || toneMap != o.toneMap || toneNorm != o.toneNorm ||
        gamut != o.gamut || sdrWhite != o.sdrWhite || hdrPeak != o.hdrPeak || cinemaPeak != o.cinemaPeak ||
        hdrRefWhite != o.hdrRefWhite || cinemaRefWhite != o.cinemaRefWhite || knee != o.knee ||
        gamutThreshold != o.gamutThreshold || lut3dSize != o.lut3dSize)
        return false;
    if (lut3d == o.lut3d) return true;
    if (!lut3d || !o.lut3d) return false;
    return *lut3d == *o.lut3d;
}

namespace {

// SMPTE ST 2084 (PQ) and BT.2100 HLG in double, for the tables.
constexpr double kM1 = 2610.0 / 16384.0, kM2 = 2523.0 / 4096.0 * 128.0, kC1 = 3424.0 / 4096.0,
                 kC2 = 2413.0 / 4096.0 * 32.0, kC3 = 2392.0 / 4096.0 * 32.0;

double pqToNits(double e) {
    e = std::clamp(e, 0.0, 1.0);
    const double p = std::pow(e, 1.0 / kM2);
    return 10000.0 * std::pow(std::max(p - kC1, 0.0) / (kC2 - kC3 * p), 1.0 / kM1);
}

double nitsToPq(double nits) {
    const double y = std::pow(std::clamp(nits / 10000.0, 0.0, 1.0), kM1);
    return std::pow((kC1 + kC2 * y) / (1.0 + kC3 * y), kM2);
}

double hlgSceneLight(double e) {
    const double a = 0.17883277, b = 1.0 - 4.0 * a, c = 0.5 - a * std::log(4.0 * a);
    return e <= 0.5 ? e * e / 3.0 : (std::exp((e - c) / a) + b) / 12.0;
}

}  // namespace

void buildColorTables(const ColorSpec& c, std::vector<float>& lut, float cp[kColorParams]) {
    const bool hdr = c.transfer == 2 || c.transfer == 3, passthrough = c.primaries == 4;
    lut.resize(65536);
    for (int i = 0; i < 65536; ++i) {
        const double e = i / 65535.0;
        double nits;
        switch (c.transfer) {
        case 1: nits = (e <= 0.04045 ? e / 12.92 : std::pow((e + 0.055) / 1.055, 2.4)) * c.sdrWhite; break;
        case 2: nits = c.toneMap == 0 ? std::pow(e, 2.4) * c.sdrWhite : pqToNits(e); break;
        case 3: nits = c.toneMap == 0 ? std::pow(e, 2.4) * c.sdrWhite : 1000.0 * std::pow(hlgSceneLight(e), 1.2); break;
        case 4: nits = e * c.sdrWhite; break;
        case 5: nits = std::pow(e, 2.6) * 52.37; break;
        default: nits = std::pow(e, 2.4) * c.sdrWhite; break;
        }
        lut[i] = (float)nits;
    }
    std::fill(cp, cp + kColorParams, 0.f);
    double M[9];
    rgbToXyz(c.primaries, M);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) cp[4 * r + k] = (float)M[3 * r + k];
    cp[17] = (float)c.cinemaPeak;
    if (hdr && !passthrough && c.toneMap == 1) {  // shoulder (scene-relative)
        const double gain = c.cinemaRefWhite / c.cinemaPeak;
        double knee = std::clamp(c.knee, 0.05, 0.99);
        const double vmax = gain * c.hdrPeak / c.hdrRefWhite;  // source peak on the output scale
        if (vmax <= 1.0) knee = 1.0;                            // it already fits: no shoulder
        cp[12] = 1.f;
        cp[13] = (float)c.hdrRefWhite;
        cp[14] = (float)gain;
        cp[15] = (float)knee;
        cp[16] = (float)std::max(1e-3, (vmax - knee) / (1.0 - knee));
    } else if (hdr && !passthrough && c.toneMap == 2 && c.hdrPeak > c.cinemaPeak) {  // BT.2390 EETF
        cp[12] = 2.f;
        cp[39] = (float)nitsToPq(c.hdrPeak);
        cp[40] = (float)(nitsToPq(c.cinemaPeak) / nitsToPq(c.hdrPeak));
        cp[41] = std::max(0.f, 1.5f * cp[40] - 0.5f);
    }
    cp[42] = c.toneNorm == 0 ? 2.f : c.toneNorm == 1 ? 1.f : 0.f;
    if (!passthrough) {
        double T[9], Ti[9];
        rgbToXyz(c.gamut == 1 ? 3 : 2, T);
        invert3(T, Ti);
        for (int i = 0; i < 9; ++i) {
            cp[20 + i] = (float)(Ti[i] / c.cinemaPeak);
            cp[29 + i] = (float)(T[i] * c.cinemaPeak);
        }
        if (c.gamut != 2) {
            cp[18] = 1.f;
            cp[19] = (float)std::clamp(c.gamutThreshold, 0.0, 0.99);
        }
    }
    cp[38] = (float)(c.lut3d && c.lut3dSize > 1 ? c.lut3dSize : 0);
}

// A 3D LUT in the .cube text format (Adobe Cube LUT Specification 1.0): keyword lines
// (TITLE, LUT_3D_SIZE, DOMAIN_MIN / DOMAIN_MAX) ahead of size^3 data lines of three numbers
// each, red varying fastest; '#' starts a comment; blank lines and keywords this reader does
// not use are skipped. Only the unit domain is accepted, and 1D LUTs are not.
std::vector<float> readCubeLut(const std::string& path, int& size) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error(".cube LUT " + path + ": cannot be opened");
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    auto bad = [&](int lineNo, const std::string& why) {
        return std::runtime_error(".cube LUT " + path + ", line " + std::to_string(lineNo) + ": " + why);
    };
    // up to `want` numbers from the text at p (strtof), how many were read; *end past them
    auto numbers = [](const char* p, float* out, int want, const char** end) {
        int n = 0;
        for (; n < want; ++n) {
            char* e;
            const float v = std::strtof(p, &e);
            if (e == p) break;
            out[n] = v;
            p = e;
        }
        *end = p;
        return n;
    };
    auto blank = [](const char* p) {
        while (*p == ' ' || *p == '\t') ++p;
        return *p == 0;
    };
    std::vector<float> data;
    size = 0;
    int lineNo = 0;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        ++lineNo;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        const size_t first = line.find_first_not_of(" \t");
        if (first == std::string::npos) continue;
        const char* p = line.c_str() + first;
        if (std::isalpha((unsigned char)*p)) {  // a keyword line
            const char* q = p;
            while (*q && *q != ' ' && *q != '\t') ++q;
            const std::string key(p, q);
            if (!data.empty()) throw bad(lineNo, "keyword " + key + " after the table");
            float v[3];
            const char* end;
            if (key == "LUT_3D_SIZE") {
                if (numbers(q, v, 1, &end) != 1 || !blank(end) || v[0] != std::floor(v[0]))
                    throw bad(lineNo, "LUT_3D_SIZE needs one whole number");
                size = (int)v[0];
                if (size < 2 || size > 256) throw bad(lineNo, "LUT_3D_SIZE must be 2 to 256");
            } else if (key == "LUT_1D_SIZE") {
                throw bad(lineNo, "1D LUTs are not supported");
            } else if (key == "DOMAIN_MIN" || key == "DOMAIN_MAX") {
                const float unit = key == "DOMAIN_MIN" ? 0.f : 1.f;
                if (numbers(q, v, 3, &end) != 3 || !blank(end)) throw bad(lineNo, key + " needs three numbers");
                if (v[0] != unit || v[1] != unit || v[2] != unit) throw bad(lineNo, "only the domain 0..1 is supported");
            }
            continue;  // TITLE and keywords of other tools
        }
        float v[3];
        const char* end;
        if (numbers(p, v, 3, &end) != 3 || !blank(end)) throw bad(lineNo, "a table line needs exactly three numbers");
        data.insert(data.end(), v, v + 3);
    }
    if (size == 0) throw std::runtime_error(".cube LUT " + path + ": no LUT_3D_SIZE");
    const size_t want = (size_t)size * size * size * 3;
    if (data.size() != want)
        throw std::runtime_error(".cube LUT " + path + ": " + std::to_string(data.size() / 3) + " table lines, " +
                                 std::to_string(want / 3) + " expected");
    return data;
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
