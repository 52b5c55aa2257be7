// C entry points over the reference's own CPU hot-path sources (layout.cpp,
// dwt_cpu.cpp, quant.cpp, t1_ref.cpp, color.cpp), compiled from where they lie by
// oracle/Makefile into oracle/_ref/libgailref.so. TEST INFRASTRUCTURE ONLY: used to
// write golden vectors (tests/golden/make_golden.py), to cross-check the oracle, and
// as the timed CPU baseline (the frame flow of cpu_backend.cpp:61-132 over the
// reference's functions, on a plain thread pool of this file's own).
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "core/color.h"
#include "core/dwt_cpu.h"
#include "core/layout.h"
#include "core/quant.h"
#include "core/t1_ref.h"

using namespace gail;

namespace {

// Same field order as or_band in oracle.h.
struct RefBand {
    int32_t comp, res, level, orient, px0, py0, w, h, ncbx, ncby, cblk_base, eps, mu, mb;
    float delta, pad;
    double weight;
};
struct RefCblk {
    int32_t band, x0, y0, w, h;
};
struct RefT1 {
    int32_t nbps, num_passes;
    uint32_t len;
    uint32_t pass_len[64];
    int32_t pass_dist[64];
};
struct RefInfo {
    uint32_t nbps, num_passes, data_offset, data_len, pass_base;
};

void parallelFor(int threads, int n, const std::function<void(int)>& f) {
    std::atomic<int> next{0};
    auto work = [&] {
        for (int i; (i = next.fetch_add(1)) < n;) f(i);
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < threads; ++t) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
}

ColorParams colorParams(int transfer, int primaries, double sdrWhite, double cinemaPeak) {
    ColorParams p;
    p.transfer = (Transfer)transfer;
    p.primaries = (Primaries)primaries;
    p.sdrWhite = sdrWhite;
    p.cinemaPeak = cinemaPeak;
    return p;
}

}  // namespace

extern "C" {

int ref_layout(int W, int H, int fourk, double baseStep, RefBand* bands, int* nbands, RefCblk* cblks, int cap) {
    try {
        const FrameLayout L = makeLayout(W, H, fourk ? Profile::Cinema4K : Profile::Cinema2K, baseStep);
        *nbands = (int)L.bands.size();
        for (size_t i = 0; i < L.bands.size(); ++i) {
            const Subband& b = L.bands[i];
            bands[i] = {b.comp, b.res, b.level, b.orient, b.px0, b.py0, b.w, b.h, b.ncbx, b.ncby, b.cblkBase,
                        b.eps, b.mu, b.Mb, b.delta, 0.f, b.weight};
        }
        for (int i = 0; i < L.numCblks() && i < cap; ++i) {
            const CodeBlock& c = L.cblks[i];
            cblks[i] = {c.band, c.x0, c.y0, c.w, c.h};
        }
        return L.numCblks();
    } catch (...) {
        return -1;
    }
}

double ref_synthesis_norm(int level, int orient) { return synthesisNorm(level, orient); }

void ref_color_tables(int transfer, int primaries, double sdrWhite, double cinemaPeak, float* lut, float* params) {
    const ColorConverter cc(colorParams(transfer, primaries, sdrWhite, cinemaPeak));
    std::memcpy(lut, cc.lut().data(), 65536 * sizeof(float));
    std::memcpy(params, cc.params().data(), cc.params().size() * sizeof(float));
}

// The managed route with every colour parameter (or_color_cfg's layout), and an app 3D
// LUT when lut3dSize > 1.
struct RefColorCfg {
    int transfer, primaries, toneMap, toneNorm, gamut, lut3dSize;
    double sdrWhite, hdrPeak, cinemaPeak, hdrRefWhite, cinemaRefWhite, knee, gamutThreshold;
};

static ColorParams colorParamsFull(const RefColorCfg& c, const float* lut3) {
    ColorParams p = colorParams(c.transfer, c.primaries, c.sdrWhite, c.cinemaPeak);
    p.toneMap = (ToneMap)c.toneMap;
    p.toneNorm = (ToneNorm)c.toneNorm;
    p.gamut = (TargetGamut)c.gamut;
    p.hdrPeak = c.hdrPeak;
    p.hdrRefWhite = c.hdrRefWhite;
    p.cinemaRefWhite = c.cinemaRefWhite;
    p.knee = c.knee;
    p.gamutThreshold = c.gamutThreshold;
    if (c.lut3dSize > 1 && lut3) {
        p.lut3dData.assign(lut3, lut3 + (size_t)c.lut3dSize * c.lut3dSize * c.lut3dSize * 3);
        p.lut3dSize = c.lut3dSize;
    }
    return p;
}

int ref_color_setup(const RefColorCfg* c, const float* lut3, float* lut, float* params) {
    try {
        const ColorConverter cc(colorParamsFull(*c, lut3));
        std::memcpy(lut, cc.lut().data(), 65536 * sizeof(float));
        std::memcpy(params, cc.params().data(), cc.params().size() * sizeof(float));
        return 0;
    } catch (...) {
        return -1;
    }
}

int ref_color_pixels(const RefColorCfg* c, const float* lut3, int n, const uint16_t* rgb, int* out) {
    try {
        const ColorConverter cc(colorParamsFull(*c, lut3));
        for (int i = 0; i < n; ++i) cc.toXYZ12(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], out + 3 * i);
        return 0;
    } catch (...) {
        return -1;
    }
}

void ref_convert_frame(int transfer, int primaries, double sdrWhite, double cinemaPeak, const uint16_t* rgb, int W,
                       int H, float* planes, uint16_t* xyz12) {
    const ColorConverter cc(colorParams(transfer, primaries, sdrWhite, cinemaPeak));
    float* pl[3] = {planes, planes + (size_t)W * H, planes + 2 * (size_t)W * H};
    cc.convertFrame(rgb, W, H, pl, xyz12);
}

void ref_planar_to_rgb48(const void* src, int bits, int aw, int ah, int sx, int sy, int rgb, int fullRange,
                         int chroma, double kr, double kb, int W, int H, int ox, int oy, uint16_t* out) {
    FrameFormat f;
    f.kind = FrameFormat::Kind::Planar;
    f.chroma = (FrameFormat::Chroma)chroma;
    f.width = W;
    f.height = H;
    f.activeW = aw;
    f.activeH = ah;
    f.ox = ox;
    f.oy = oy;
    f.sx = sx;
    f.sy = sy;
    f.rgb = rgb != 0;
    f.bits = bits;
    f.fullRange = fullRange != 0;
    f.kr = kr;
    f.kb = kb;
    planarToRgb48(f, src, out);
}

void ref_fdwt97_2d(float* plane, int W, int H, int levels) { fdwt97_2d(plane, W, H, levels); }

void ref_fdwt97_1d_f64(const double* in, double* out, int n) { fdwt97_1d<double>(in, out, n); }

uint32_t ref_quantize_coeff(float v, float invDelta, uint32_t maxMag) { return quantizeCoeff(v, invDelta, maxMag); }

int ref_t1_encode(const uint32_t* coeffs, int w, int h, int orient, int maxPasses, float slopeThresh, float distScale,
                  uint8_t* out, uint32_t outCap, RefT1* res, uint8_t* syms, uint32_t symCap, uint32_t* nsyms,
                  uint32_t* passEnd) {
    try {
        T1Result r;
        T1Trace tr;
        const T1EarlyStop stop{slopeThresh, distScale};
        t1EncodeRef(coeffs, w, h, orient, r, syms ? &tr : nullptr, maxPasses, &stop);
        std::memset(res, 0, sizeof *res);
        res->nbps = r.nbps;
        res->num_passes = r.numPasses;
        res->len = (uint32_t)r.data.size();
        if (r.data.size() > outCap) return -1;
        if (!r.data.empty()) std::memcpy(out, r.data.data(), r.data.size());
        for (int k = 0; k < r.numPasses; ++k) {
            res->pass_len[k] = r.passLen[k];
            res->pass_dist[k] = r.passDist[k];
        }
        if (syms) {
            *nsyms = (uint32_t)tr.symbols.size();
            std::memcpy(syms, tr.symbols.data(), std::min<size_t>(symCap, tr.symbols.size()));
            for (size_t k = 0; k < tr.passEnd.size(); ++k) passEnd[k] = tr.passEnd[k];
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// One frame of planar 12-bit X'Y'Z' through the reference's CPU flow (cpu_backend.cpp:61-132):
// fill() widens the codes to 16 bits (gail_api.cpp:389-394), ColorConverter maps them back
// (pass-through) with DC shift and ICT, fdwt97_2d, then quantizeCodeBlock + t1EncodeRef per
// block. Returns the total codeword bytes (-1 on error).
int64_t ref_encode_frame_xyz(const uint16_t* xyz, int W, int H, int fourk, double baseStep, float slopeThresh,
                             int threads, RefInfo* info, uint32_t* passLen, int32_t* passDist, uint8_t* data,
                             uint64_t dataCap, uint32_t* coeffsOut, float* planesOut) {
    try {
        const FrameLayout L = makeLayout(W, H, fourk ? Profile::Cinema4K : Profile::Cinema2K, baseStep);
        ColorParams cp;
        cp.transfer = Transfer::Gamma26;
        cp.primaries = Primaries::XYZ;
        const ColorConverter cc(cp);
        const size_t npix = (size_t)W * H;
        std::vector<float> own;
        float* base = planesOut;
        if (!base) {
            own.resize(3 * npix);
            base = own.data();
        }
        float* pl[3] = {base, base + npix, base + 2 * npix};
        parallelFor(threads, H, [&](int y) {
            std::vector<uint16_t> row((size_t)W * 3);
            for (int x = 0; x < W; ++x)
                for (int c = 0; c < 3; ++c) {
                    const uint32_t v = std::min<uint32_t>(xyz[c * npix + (size_t)y * W + x], 4095u);
                    row[3 * x + c] = (uint16_t)((v * 65535u + 2047u) / 4095u);
                }
            float* rows[3] = {pl[0] + (size_t)y * W, pl[1] + (size_t)y * W, pl[2] + (size_t)y * W};
            cc.convertFrame(row.data(), W, 1, rows, nullptr);
        });
        const ParallelFor par = [&](int n, const std::function<void(int)>& f) { parallelFor(threads, n, f); };
        fdwt97_2d(pl, 3, W, H, L.levels, &par);
        const int n = L.numCblks();
        std::vector<T1Result> res(n);
        std::vector<uint32_t> ownCoeffs;
        uint32_t* coeffs = coeffsOut;
        if (!coeffs) {
            ownCoeffs.resize((size_t)n * kCblkArea);
            coeffs = ownCoeffs.data();
        }
        const float* cpl[3] = {pl[0], pl[1], pl[2]};
        std::vector<float> scale;
        for (const Subband& b : L.bands) scale.push_back((float)((double)b.delta * b.delta * b.weight));
        parallelFor(threads, n, [&](int i) {
            const CodeBlock& cb = L.cblks[i];
            uint32_t* q = coeffs + (size_t)i * kCblkArea;
            quantizeCodeBlock(L, cpl, i, q);
            const T1EarlyStop stop{slopeThresh, scale[cb.band]};
            t1EncodeRef(q, cb.w, cb.h, L.bands[cb.band].orient, res[i], nullptr, kMaxPasses, &stop);
        });
        uint64_t off = 0;
        uint32_t pb = 0;
        for (int i = 0; i < n; ++i) {
            const T1Result& r = res[i];
            if (off + r.data.size() > dataCap) return -1;
            info[i] = {(uint32_t)r.nbps, (uint32_t)r.numPasses, (uint32_t)off, (uint32_t)r.data.size(), pb};
            if (!r.data.empty()) std::memcpy(data + off, r.data.data(), r.data.size());
            for (int k = 0; k < r.numPasses; ++k) {
                passLen[pb + k] = r.passLen[k];
                passDist[pb + k] = r.passDist[k];
            }
            pb += r.numPasses;
            off += r.data.size();
        }
        return (int64_t)off;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
