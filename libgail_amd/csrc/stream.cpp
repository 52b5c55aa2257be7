// Per-stream set-up shared by the encoder and the server of libgail.so (see stream.h).
#include "stream.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstring>
#include <mutex>

namespace gail_amd {

std::string& lastErrorText() {
    thread_local std::string text;
    return text;
}

const PluginHandle& loadPlugin() {
    static PluginHandle h;
    static std::once_flag once;
    std::call_once(once, [] {
        Dl_info info;
        if (!dladdr((void*)&loadPlugin, &info) || !info.dli_fname) {
            h.error = "cannot locate libgail.so";
            return;
        }
        std::string dir = info.dli_fname;
        dir = dir.substr(0, dir.find_last_of('/') + 1);
        void* lib = dlopen((dir + "libgail_hip.so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!lib) {
            h.error = std::string("HIP backend plugin not loaded: ") + dlerror();
            return;
        }
        auto get = (const GailHipPlugin* (*)(uint32_t))dlsym(lib, GAIL_HIP_PLUGIN_ENTRY);
        if (!get) {
            h.error = "libgail_hip.so has no " GAIL_HIP_PLUGIN_ENTRY " entry point";
            return;
        }
        const GailHipPlugin* t = get(GAIL_HIP_PLUGIN_ABI);
        // the table is read only once abi and size say it has the fields read
        if (!t) h.error = "libgail_hip.so does not support plugin ABI " + std::to_string(GAIL_HIP_PLUGIN_ABI);
        else if (t->abi != GAIL_HIP_PLUGIN_ABI)
            h.error = "libgail_hip.so is from another build (plugin ABI " + std::to_string(t->abi) + ", this library " +
                      std::to_string(GAIL_HIP_PLUGIN_ABI) + ")";
        else if (t->size != sizeof(GailHipPlugin))
            h.error = "libgail_hip.so is from another build (plugin table of " + std::to_string(t->size) + " bytes, " +
                      std::to_string(sizeof(GailHipPlugin)) + " expected)";
        else if (t->layout != gail_hip_plugin_layout())
            h.error = "libgail_hip.so is from another build (its structure layout differs)";
        else
            h.table = t;
    });
    return h;
}

bool allZero(const uint64_t* p, int n) {
    for (int i = 0; i < n; ++i)
        if (p[i]) return false;
    return true;
}

uint64_t memoryHeadroom(uint64_t total) {
    const uint64_t floor = 512ull << 20, share = total / 100 * 6 + total % 100 * 6 / 100;
    return share > floor ? share : floor;
}

int fitFrames(uint64_t available, uint64_t total, uint64_t perFrame, int requested, uint64_t limit) {
    const int want = requested > 1 ? requested : 1;
    if (!perFrame || (!available && !limit)) return want;
    uint64_t room = limit;
    if (available) {
        const uint64_t head = memoryHeadroom(total), usable = available > head ? available - head : 0;
        room = limit && limit < usable ? limit : usable;
    }
    const uint64_t fit = room / (perFrame + perFrame / 10);
    return fit < 1 ? 1 : (fit < (uint64_t)want ? (int)fit : want);
}

bool planarInfo(gail_pixel_format f, PlanarInfo& p) {
    // the planar formats of gail_pixel_format, in enumeration order from GAIL_PIX_YUV420P:
    // three Y'CbCr depths each at 4:2:0, 4:2:2 and 4:4:4, then G'B'R' at 8, 10, 12 and 16 bits
    static const PlanarInfo table[] = {
        {8, 1, 1, 0}, {10, 1, 1, 0}, {12, 1, 1, 0}, {8, 1, 0, 0}, {10, 1, 0, 0}, {12, 1, 0, 0}, {8, 0, 0, 0},
        {10, 0, 0, 0}, {12, 0, 0, 0}, {8, 0, 0, 1}, {10, 0, 0, 1}, {12, 0, 0, 1}, {16, 0, 0, 1}};
    const int i = (int)f - (int)GAIL_PIX_YUV420P;
    if (i < 0 || i >= (int)(sizeof table / sizeof table[0])) return false;
    p = table[i];
    return true;
}

std::unique_ptr<StreamSpec> StreamSpec::make(const gail_config& c) {
    std::unique_ptr<StreamSpec> e(new StreamSpec());
    gail_config_init(&e->cfg, sizeof e->cfg);
    std::memcpy(&e->cfg, &c, std::min(c.struct_size ? c.struct_size : sizeof c, sizeof e->cfg));
    gail_config& cfg = e->cfg;
    if (cfg.backend != GAIL_BACKEND_AUTO && cfg.backend != GAIL_BACKEND_HIP)
        throw ApiError(GAIL_ERROR_UNSUPPORTED, "this library has the HIP backend only");
    if (!(cfg.fps > 0) || !(cfg.max_mbps > 0) || cfg.max_mbps > 250.0)
        throw ApiError(GAIL_ERROR_INVALID, "fps or max_mbps out of range");
    if (!allZero(cfg.input.reserved, 4) || !allZero(cfg.color.reserved, 8))
        throw ApiError(GAIL_ERROR_INVALID, "reserved fields must be zero");
    if (cfg.color.mode != GAIL_COLOR_MANAGED && cfg.color.mode != GAIL_COLOR_LUT3D)
        throw ApiError(GAIL_ERROR_INVALID, "invalid color mode");
    if (gail_container_size(cfg.profile, cfg.container, &e->width, &e->height) != GAIL_OK)
        throw ApiError(GAIL_ERROR_INVALID, "invalid profile or container");
    e->picW = cfg.input.width ? cfg.input.width : e->width;
    e->picH = cfg.input.height ? cfg.input.height : e->height;
    if (e->picW <= 0 || e->picH <= 0 || e->picW > e->width || e->picH > e->height)
        throw ApiError(GAIL_ERROR_INVALID, "picture larger than the container");
    e->ox = (e->width - e->picW) / 2;
    e->oy = (e->height - e->picH) / 2;
    const bool fourk = cfg.profile == GAIL_PROFILE_4K;
    e->L = makeLayout(e->width, e->height, fourk, 1.0);
    for (const Subband& b : e->L.bands) {
        GailHipSubband s{};
        s.comp = b.comp;
        s.res = b.res;
        s.level = b.level;
        s.orient = b.orient;
        s.px0 = b.px0;
        s.py0 = b.py0;
        s.w = b.w;
        s.h = b.h;
        s.ncbx = b.ncbx;
        s.ncby = b.ncby;
        s.cblkBase = b.cblkBase;
        s.eps = b.eps;
        s.mu = b.mu;
        s.Mb = b.Mb;
        s.delta = b.delta;
        s.weight = b.weight;
        e->bands.push_back(s);
    }
    e->plainLayout = GailHipLayout{fourk ? 1 : 0, e->width, e->height, e->L.levels, e->L.guardBits,
                                   (int)e->bands.size(), e->L.numCblks(), 1.0, e->bands.data()};
    GailHipEncodeSettings& s = e->settings;
    s.profile = fourk ? 1 : 0;
    s.width = e->width;
    s.height = e->height;
    s.baseStep = 1.0;
    s.slopeFactor = e->slopeFactor;
    s.bitsPerSymbol = 0.5;
    s.input.width = e->width;
    s.input.height = e->height;
    PlanarInfo pi;
    const gail_pixel_format fmt = cfg.input.format;
    if (planarInfo(fmt, pi)) {
        s.input.kind = 1;
        s.input.activeW = e->picW;
        s.input.activeH = e->picH;
        s.input.ox = e->ox;
        s.input.oy = e->oy;
        s.input.sx = pi.sx;
        s.input.sy = pi.sy;
        s.input.rgb = pi.rgb;
        s.input.bits = pi.bits;
        s.input.fullRange = pi.rgb ? 1 : (cfg.input.full_range ? 1 : 0);
        s.input.chroma = (int)cfg.input.chroma_location;
        const double kr[3] = {0.2126, 0.2627, 0.299}, kb[3] = {0.0722, 0.0593, 0.114};
        if (cfg.input.matrix < GAIL_MATRIX_BT709 || cfg.input.matrix > GAIL_MATRIX_BT601)
            throw ApiError(GAIL_ERROR_INVALID, "invalid matrix");
        s.input.kr = kr[cfg.input.matrix];
        s.input.kb = kb[cfg.input.matrix];
        // chroma planes of ceil(w / 2^sx) x ceil(h / 2^sy) samples, one or two bytes each
        const size_t bytesPer = pi.bits > 8 ? 2 : 1;
        const int cw = (e->picW + pi.sx) >> pi.sx, ch = (e->picH + pi.sy) >> pi.sy;
        e->planar = true;
        size_t offset = 0;
        for (int c = 0; c < 3; ++c) {
            e->planes[c].offset = offset;
            e->planes[c].rowBytes = (size_t)(c ? cw : e->picW) * bytesPer;
            e->planes[c].rows = c ? ch : e->picH;
            offset += e->planes[c].rowBytes * e->planes[c].rows;
        }
        e->inBytes = offset;
    } else if (fmt == GAIL_PIX_RGB48) {
        s.input.kind = 0;
        e->inBytes = (size_t)e->width * e->height * 6;
    } else if (fmt == GAIL_PIX_XYZ12 || fmt == GAIL_PIX_XYZ12_MSB || fmt == GAIL_PIX_XYZ12_PLANAR) {
        s.input.kind = 2;
        e->inBytes = (size_t)e->width * e->height * 6;
    } else {
        throw ApiError(GAIL_ERROR_INVALID, "invalid pixel format");
    }
    // colour (gail_api.cpp:158-198): gail_color in ColorParams' terms
    const gail_color& gc = cfg.color;
    if (!(gc.sdr_white > 0 && gc.hdr_peak > 0 && gc.hdr_ref_white > 0 && gc.cinema_ref_white > 0 &&
          gc.cinema_peak > 0 && gc.cinema_peak <= 52.37 && gc.knee > 0 && gc.knee < 1 && gc.gamut_threshold >= 0 &&
          gc.gamut_threshold < 1))
        throw ApiError(GAIL_ERROR_INVALID, "color parameter out of range");
    if (gc.transfer < GAIL_TRANSFER_BT1886 || gc.transfer > GAIL_TRANSFER_GAMMA26 || gc.primaries < GAIL_PRIMARIES_BT709 ||
        gc.primaries > GAIL_PRIMARIES_DCIP3 || gc.tone_map < GAIL_TONE_MAP_SIGNAL || gc.tone_map > GAIL_TONE_MAP_NONE ||
        gc.tone_norm < GAIL_TONE_NORM_CHANNEL || gc.tone_norm > GAIL_TONE_NORM_LUMINANCE || gc.gamut < GAIL_GAMUT_NONE ||
        gc.gamut > GAIL_GAMUT_DCIP3)
        throw ApiError(GAIL_ERROR_INVALID, "invalid colour enumeration");
    const bool xyz = s.input.kind == 2;
    GailHipColor& col = s.color;
    col.transfer = xyz ? 5 : (int)gc.transfer;
    col.primaries = xyz ? 4 : (int)gc.primaries;
    col.toneMap = (int)gc.tone_map;
    col.toneNorm = (int)gc.tone_norm;
    col.gamut = gc.gamut == GAIL_GAMUT_P3D65 ? 0 : gc.gamut == GAIL_GAMUT_DCIP3 ? 1 : 2;
    col.sdrWhite = gc.sdr_white;
    col.hdrPeak = gc.hdr_peak;
    col.cinemaPeak = gc.cinema_peak;
    col.hdrRefWhite = gc.hdr_ref_white;
    col.cinemaRefWhite = gc.cinema_ref_white;
    col.knee = gc.knee;
    col.gamutThreshold = gc.gamut_threshold;
    if (gc.mode == GAIL_COLOR_LUT3D && !xyz) {
        if (!gc.lut3d || gc.lut3d_size < 2 || gc.lut3d_size > 256) throw ApiError(GAIL_ERROR_INVALID, "invalid 3D LUT");
        const size_t n = (size_t)gc.lut3d_size * gc.lut3d_size * gc.lut3d_size * 3;
        e->lut3d.assign(gc.lut3d, gc.lut3d + n);
        col.lut3dSize = gc.lut3d_size;
        col.lut3dData = e->lut3d.data();
        col.lut3dCount = n;
    }
    e->limits = dciLimits(e->eyeFps(), cfg.max_mbps, 200.0);
    if (emptyCodestreamBytes(e->L) > e->limits.maxTotal)
        throw ApiError(GAIL_ERROR_INVALID, "the frame budget is below the codestream headers");
    return e;
}

int StreamSpec::checkFrame(const gail_frame& f, std::string& why) const {
    if (!allZero(f.reserved, 4)) {
        why = "gail_frame.reserved must be zero";
        return GAIL_ERROR_INVALID;
    }
    PlanarInfo pi;
    const int planes = (planarInfo(cfg.input.format, pi) || cfg.input.format == GAIL_PIX_XYZ12_PLANAR) ? 3 : 1;
    for (int c = 0; c < planes; ++c)
        if (!f.data[c]) {
            why = "frame plane missing";
            return GAIL_ERROR_INVALID;
        }
    return GAIL_OK;
}

void StreamSpec::fill(const gail_frame& f, uint8_t* dst, int part, int parts) const {
    const gail_pixel_format fmt = cfg.input.format;
    auto rowsOf = [&](int n, int& y0, int& y1) {
        y0 = (int)((long long)n * part / parts);
        y1 = (int)((long long)n * (part + 1) / parts);
    };
    if (planar) {
        for (int c = 0; c < 3; ++c) {
            const Plane& pl = planes[c];
            const ptrdiff_t stride = f.stride[c] ? f.stride[c] : (ptrdiff_t)pl.rowBytes;
            int y0, y1;
            rowsOf(pl.rows, y0, y1);
            const uint8_t* src = static_cast<const uint8_t*>(f.data[c]) + y0 * stride;
            uint8_t* out = dst + pl.offset + (size_t)y0 * pl.rowBytes;
            if (stride == (ptrdiff_t)pl.rowBytes) {  // unpadded rows: the part in one copy
                std::memcpy(out, src, pl.rowBytes * (size_t)(y1 - y0));
                continue;
            }
            for (int y = y0; y < y1; ++y, out += pl.rowBytes, src += stride) std::memcpy(out, src, pl.rowBytes);
        }
        return;
    }
    int y0, y1;
    rowsOf(picH, y0, y1);
    if (fmt == GAIL_PIX_RGB48) {
        if (picW != width || picH != height) std::memset(dst, 0, inBytes);  // parts == 1 here
        const size_t row = (size_t)picW * 6;
        const ptrdiff_t stride = f.stride[0] ? f.stride[0] : (ptrdiff_t)row;
        for (int y = y0; y < y1; ++y)
            std::memcpy(dst + ((size_t)(y + oy) * width + ox) * 6, (const uint8_t*)f.data[0] + y * stride, row);
        return;
    }
    // X'Y'Z': planar 12-bit codes, the picture centred, code 0 around it
    if (picW != width || picH != height) std::memset(dst, 0, inBytes);  // parts == 1 here
    uint16_t* planes[3] = {(uint16_t*)dst, (uint16_t*)dst + (size_t)width * height, (uint16_t*)dst + 2 * (size_t)width * height};
    for (int y = y0; y < y1; ++y) {
        const size_t o = (size_t)(y + oy) * width + ox;
        if (fmt == GAIL_PIX_XYZ12_PLANAR) {
            // the rows as they are: the device reads codes above 4095 as 4095 (the DWT's DC
            // shift clips on load), as it does for frames the served form takes directly
            for (int c = 0; c < 3; ++c) {
                const ptrdiff_t stride = f.stride[c] ? f.stride[c] : (ptrdiff_t)picW * 2;
                std::memcpy(planes[c] + o, (const uint8_t*)f.data[c] + y * stride, (size_t)picW * 2);
            }
            continue;
        }
        const ptrdiff_t stride = f.stride[0] ? f.stride[0] : (ptrdiff_t)picW * 6;
        const uint16_t* s = (const uint16_t*)((const uint8_t*)f.data[0] + y * stride);
        const bool msb = fmt == GAIL_PIX_XYZ12_MSB;
        for (int x = 0; x < picW; ++x)
            for (int c = 0; c < 3; ++c) {
                const uint32_t v = msb ? s[3 * x + c] >> 4 : s[3 * x + c];
                planes[c][o + x] = (uint16_t)std::min<uint32_t>(v, 4095u);
            }
    }
}

const void* StreamSpec::directInput(const gail_frame& f) const {
    const gail_pixel_format fmt = cfg.input.format;
    const uint8_t* base = static_cast<const uint8_t*>(f.data[0]);
    if (!base) return nullptr;
    if (planar) {
        for (int c = 0; c < 3; ++c) {
            if (f.data[c] != base + planes[c].offset) return nullptr;
            if (f.stride[c] && (size_t)f.stride[c] != planes[c].rowBytes) return nullptr;
        }
        return base;
    }
    if (picW != width || picH != height) return nullptr;
    if (fmt == GAIL_PIX_RGB48) return (!f.stride[0] || (size_t)f.stride[0] == (size_t)width * 6) ? base : nullptr;
    if (fmt != GAIL_PIX_XYZ12_PLANAR) return nullptr;  // interleaved X'Y'Z' is rearranged on the host
    const size_t plane = (size_t)width * height * 2;
    for (int c = 0; c < 3; ++c) {
        if (f.data[c] != base + c * plane) return nullptr;
        if (f.stride[c] && (size_t)f.stride[c] != (size_t)width * 2) return nullptr;
    }
    return base;
}

}  // namespace gail_amd
