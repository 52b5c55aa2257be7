// C entry points over the reference's own CPU sources (layout.cpp, dwt_cpu.cpp, quant.cpp,
// t1_ref.cpp, color.cpp; for decoding t1_dec.cpp, cpu_decode_backend.cpp,
// codestream_parse.cpp, t2_parse.cpp; the DCI checker codestream_check.cpp), compiled from where they lie by
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

#include "core/codestream_check.h"
#include "core/codestream_parse.h"
#include "core/color.h"
#include "core/decode_backend.h"
#include "core/dwt_cpu.h"
#include "core/layout.h"
#include "core/quant.h"
#include "core/t1_dec.h"
#include "core/t1_ref.h"
#include "core/t2_parse.h"
#include "core/thread_pool.h"

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

// t1DecodeBlock (t1_dec.cpp:175-178), words returned with row stride w.
int ref_t1_decode_words(const uint8_t* data, uint32_t len, int w, int h, int orient, int nbps, int passes,
                        uint32_t* words) {
    if (w < 1 || h < 1 || w > kCblkSize || h > kCblkSize) return -1;
    uint32_t q[kCblkArea];
    t1DecodeBlock(data, len, w, h, orient, nbps, passes, q);
    for (int y = 0; y < h; ++y) std::memcpy(words + (size_t)y * w, q + (size_t)y * kCblkSize, sizeof(uint32_t) * w);
    return 0;
}

// idwt97_2d (dwt_cpu.cpp:213-240) of nplanes planes stored back to back.
void ref_idwt97_2d(float* planes, int nplanes, int W, int H, int levels, int skip) {
    std::vector<float*> pl;
    for (int c = 0; c < nplanes; ++c) pl.push_back(planes + (size_t)c * W * H);
    idwt97_2d(pl.data(), nplanes, W, H, levels, skip, nullptr);
}

struct RefCblkCode {
    uint32_t offset, len;
    uint8_t passes, nbps, pad[2];
};

// The reference's CPU decode slot (cpu_decode_backend.cpp) for the layout of (W, H,
// profile, baseStep): blocks[nblocks] located in data[size]; the picture goes to dst (rows
// strideBytes apart), the planes before the inverse wavelet to wavelet (may be NULL).
int ref_decode_frame(int W, int H, int fourk, double baseStep, const uint8_t* data, size_t size,
                     const RefCblkCode* blocks, int nblocks, int reduce, int highBits, int threads, uint16_t* dst,
                     ptrdiff_t strideBytes, float* wavelet) {
    try {
        const FrameLayout L = makeLayout(W, H, fourk ? Profile::Cinema4K : Profile::Cinema2K, baseStep);
        if (nblocks != L.numCblks()) return -1;
        DecodeSettings s;
        s.reduce = reduce;
        s.highBits = highBits != 0;
        auto pool = std::make_shared<ThreadPool>(threads, false);
        auto slot = createCpuDecoderSlot(L, s, pool);
        std::vector<CblkCode> codes(nblocks);
        for (int i = 0; i < nblocks; ++i) {
            codes[i].offset = blocks[i].offset;
            codes[i].len = blocks[i].len;
            codes[i].passes = blocks[i].passes;
            codes[i].nbps = blocks[i].nbps;
        }
        std::vector<float> wv[3];
        DecodeTimings tm;
        if (wavelet) tm.wavelet = wv;
        slot->decode(L, data, size, codes, dst, strideBytes, &tm);
        if (wavelet)
            for (int c = 0; c < 3; ++c) std::memcpy(wavelet + (size_t)c * W * H, wv[c].data(), sizeof(float) * W * H);
        return 0;
    } catch (...) {
        return -1;
    }
}

// The reference's header parser and Tier-2 reader (codestream_parse.cpp, t2_parse.cpp) on
// a codestream: its size, profile, signalled step sizes and every code-block's bytes.
// Returns the number of code-blocks (blocks and bands filled up to cap / OR_MAX_BANDS),
// or -1 with nothing filled when the stream is refused.
int ref_parse_blocks(const uint8_t* data, size_t size, int* W, int* H, int* fourk, RefCblkCode* blocks, int cap,
                     RefBand* bands, int* nbands) {
    try {
        StreamHeader h;
        std::string err;
        if (!parseCodestream(data, size, false, h, err) || !h.supported) return -1;
        const FrameLayout L = layoutFor(h);
        std::vector<CblkCode> out;
        if (!parsePackets(data, h, L, L.levels, out, err)) return -1;
        *W = L.width;
        *H = L.height;
        *fourk = h.rsiz == 4;
        for (int i = 0; i < (int)out.size() && i < cap; ++i)
            blocks[i] = {out[i].offset, out[i].len, out[i].passes, out[i].nbps, {0, 0}};
        *nbands = (int)L.bands.size();
        for (size_t i = 0; i < L.bands.size(); ++i) {
            const Subband& b = L.bands[i];
            bands[i] = {b.comp, b.res, b.level, b.orient, b.px0, b.py0, b.w, b.h, b.ncbx, b.ncby, b.cblkBase,
                        b.eps, b.mu, b.Mb, b.delta, 0.f, b.weight};
        }
        return (int)out.size();
    } catch (...) {
        return -1;
    }
}

// The reference's DCI compliance checker (codestream_check.cpp:43-280) on a codestream.
// Returns 1 when it passes, 0 when it does not (the errors, "; "-joined, in msg), -1 on an
// exception. facts (may be NULL) receives total bytes, the three component byte counts, the
// tile-part count, rsiz, width, height and decomposition levels.
int ref_check_codestream(const uint8_t* data, size_t size, double fps, char* msg, int cap, uint64_t* facts) {
    try {
        const CheckReport r = checkDciCodestream(data, size, fps);
        std::string all;
        for (const std::string& e : r.errors) all += (all.empty() ? "" : "; ") + e;
        if (msg && cap > 0) {
            std::strncpy(msg, all.c_str(), (size_t)cap - 1);
            msg[cap - 1] = 0;
        }
        if (facts) {
            const uint64_t f[9] = {r.totalBytes, r.compBytes[0], r.compBytes[1], r.compBytes[2], (uint64_t)r.tileParts,
                                   (uint64_t)r.rsiz, (uint64_t)r.width, (uint64_t)r.height, (uint64_t)r.levels};
            std::memcpy(facts, f, sizeof f);
        }
        return r.ok ? 1 : 0;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
