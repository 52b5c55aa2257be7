// The HIP backend on the reference library's device interfaces (see hip_host.h).
#include "hip_host.h"

#include <dlfcn.h>

#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>

#include "gail_hip_plugin.h"

namespace gail {

namespace {

struct Plugin {
    std::once_flag once;
    HipPluginInfo info;
    const GailHipPlugin* api = nullptr;
};
Plugin& plugin() {
    static Plugin p;
    return p;
}

// A failed call's status as the exception the pipeline expects of a device (backend.h:59-72):
// DeviceOutOfMemory is soft (the device stays usable), logic_error a misuse, bad_alloc host
// memory, runtime_error anything that ends the run.
void check(const GailHipPlugin* api, int status) {
    if (status == GAIL_HIP_OK) return;
    const std::string msg = std::string("hip: ") + api->last_error();
    switch (status) {
    case GAIL_HIP_DEVICE_MEMORY: throw DeviceOutOfMemory(msg);
    case GAIL_HIP_HOST_MEMORY: throw std::bad_alloc();
    case GAIL_HIP_MISUSE: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
    }
}

// A FrameLayout as the plugin takes it: the plain subband array it points to lives here.
struct PlainLayout {
    std::vector<GailHipSubband> bands;
    GailHipLayout layout{};
    explicit PlainLayout(const FrameLayout& L) {
        bands.reserve(L.bands.size());
        for (const Subband& b : L.bands) {
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
            bands.push_back(s);
        }
        layout.profile = L.profile == Profile::Cinema4K ? 1 : 0;
        layout.width = L.width;
        layout.height = L.height;
        layout.levels = L.levels;
        layout.guardBits = L.guardBits;
        layout.numBands = (int32_t)bands.size();
        layout.numCblks = L.numCblks();
        layout.baseStep = L.baseStep;
        layout.bands = bands.data();
    }
    PlainLayout(const PlainLayout&) = delete;
    PlainLayout& operator=(const PlainLayout&) = delete;
};

int transferCode(Transfer t) {
    switch (t) {
    case Transfer::BT1886: return 0;
    case Transfer::SRGB: return 1;
    case Transfer::PQ: return 2;
    case Transfer::HLG: return 3;
    case Transfer::Linear: return 4;
    case Transfer::Gamma26: return 5;
    }
    throw std::logic_error("hip: unknown transfer");
}
int primariesCode(Primaries p) {
    switch (p) {
    case Primaries::BT709: return 0;
    case Primaries::BT2020: return 1;
    case Primaries::P3D65: return 2;
    case Primaries::DCIP3: return 3;
    case Primaries::XYZ: return 4;
    }
    throw std::logic_error("hip: unknown primaries");
}

// EncodeSettings as the plugin takes them. The LUT data stays with the caller's settings
// until encoder_create has copied it.
GailHipEncodeSettings plainSettings(const EncodeSettings& s) {
    GailHipEncodeSettings o{};
    o.profile = s.profile == Profile::Cinema4K ? 1 : 0;
    o.width = s.width;
    o.height = s.height;
    o.baseStep = s.baseStep;
    o.slopeFactor = s.slopeFactor;
    o.bitsPerSymbol = s.bitsPerSymbol;
    const FrameFormat& f = s.input;
    GailHipFrameFormat& in = o.input;
    in.kind = f.kind == FrameFormat::Kind::RGB48 ? 0 : 1;
    in.chroma = f.chroma == FrameFormat::Chroma::Left ? 0 : f.chroma == FrameFormat::Chroma::Center ? 1 : 2;
    in.width = f.width;
    in.height = f.height;
    in.activeW = f.activeW;
    in.activeH = f.activeH;
    in.ox = f.ox;
    in.oy = f.oy;
    in.sx = f.sx;
    in.sy = f.sy;
    in.rgb = f.rgb ? 1 : 0;
    in.bits = f.bits;
    in.fullRange = f.fullRange ? 1 : 0;
    in.kr = f.kr;
    in.kb = f.kb;
    const ColorParams& c = s.color;
    GailHipColor& col = o.color;
    col.transfer = transferCode(c.transfer);
    col.primaries = primariesCode(c.primaries);
    col.toneMap = (int32_t)c.toneMap;    // Signal, Shoulder, BT2390, None: 0..3 in both
    col.toneNorm = (int32_t)c.toneNorm;  // Channel, MaxRGB, Luminance: 0..2 in both
    col.gamut = (int32_t)c.gamut;        // P3D65, DCIP3, None: 0..2 in both
    col.sdrWhite = c.sdrWhite;
    col.hdrPeak = c.hdrPeak;
    col.cinemaPeak = c.cinemaPeak;
    col.hdrRefWhite = c.hdrRefWhite;
    col.cinemaRefWhite = c.cinemaRefWhite;
    col.knee = c.knee;
    col.gamutThreshold = c.gamutThreshold;
    col.lut3d = c.lut3d.empty() ? nullptr : c.lut3d.c_str();
    col.lut3dSize = c.lut3dSize;
    col.lut3dData = c.lut3dData.empty() ? nullptr : c.lut3dData.data();
    col.lut3dCount = c.lut3dData.size();
    // the reference pipeline runs rate control on the host: rcMaxTotal = rcMaxComp = 0
    return o;
}

class HipEncoderSlot : public EncoderSlot {
public:
    HipEncoderSlot(const GailHipPlugin* api, GailHipEncoder* e, int width, int height)
        : api_(api), e_(e), width_(width), height_(height) {}
    ~HipEncoderSlot() override { api_->encoder_destroy(e_); }
    void* inputBuffer() override { return api_->encoder_input(e_); }
    void setSlopeThreshold(double t) override { api_->encoder_set_slope(e_, t); }
    void encode(const void* frame, Tier1Frame& out, FrameTimings* tm, FrameDebug* dbg) override {
        GailHipTier1 t{};
        GailHipFrameTimings ft{};
        GailHipDebug d{};
        d.full = dbg && dbg->full ? 1 : 0;
        check(api_, api_->encoder_encode(e_, frame, &t, tm ? &ft : nullptr, dbg ? &d : nullptr));
        static_assert(sizeof(GailHipCblkInfo) == sizeof(CblkInfo), "block record layout");
        const CblkInfo* info = reinterpret_cast<const CblkInfo*>(t.info);
        out.info.assign(info, info + t.numInfo);
        out.passLen.assign(t.passLen, t.passLen + t.numPasses);
        out.passDist.assign(t.passDist, t.passDist + t.numPasses);
        out.data.assign(t.data, t.data + t.numData);
        if (tm) {
            tm->upload = ft.upload;
            tm->color = ft.color;
            tm->dwt = ft.dwt;
            tm->quant = ft.quant;
            tm->t1 = ft.t1;
            tm->t1cm = ft.t1cm;
            tm->t1mq = ft.t1mq;
            tm->download = ft.download;
            tm->gpu = ft.gpu;
            tm->total = ft.total;
            tm->symbols = ft.symbols;
            tm->codedBytes = ft.codedBytes;
        }
        if (dbg) {
            // the plugin's X'Y'Z' tap is planar; the reference's is interleaved
            const size_t npix = (size_t)width_ * height_;
            dbg->xyz12.resize(d.xyz12 ? 3 * npix : 0);
            if (d.xyz12)
                for (size_t i = 0; i < npix; ++i)
                    for (int c = 0; c < 3; ++c) dbg->xyz12[3 * i + c] = d.xyz12[c * npix + i];
            dbg->coeffs.assign(d.coeffs, d.coeffs + (d.coeffs ? d.numCoeffs : 0));
            dbg->wavelet.assign(d.wavelet, d.wavelet + (d.wavelet ? d.numWavelet : 0));
        }
    }

private:
    const GailHipPlugin* api_;
    GailHipEncoder* e_;
    int width_, height_;
};

class HipDecoderSlot : public DecoderSlot {
public:
    HipDecoderSlot(const GailHipPlugin* api, GailHipDecoder* d) : api_(api), d_(d) {}
    ~HipDecoderSlot() override { api_->decoder_destroy(d_); }
    void decode(const FrameLayout& frame, const uint8_t* data, size_t size, const std::vector<CblkCode>& blocks,
                void* dst, ptrdiff_t strideBytes, DecodeTimings* tm) override {
        static_assert(sizeof(CblkCode) == sizeof(GailHipCblkCode), "block table record layout");
        const PlainLayout pl(frame);
        GailHipDecodeTimings t{};
        t.wantWavelet = tm && tm->wavelet ? 1 : 0;
        check(api_, api_->decoder_decode(d_, &pl.layout, data, size, reinterpret_cast<const GailHipCblkCode*>(blocks.data()),
                                         blocks.size(), dst, strideBytes, tm ? &t : nullptr));
        if (tm) {
            tm->t1 = t.t1;
            tm->dwt = t.dwt;
            tm->color = t.color;
            tm->total = t.total;
            if (tm->wavelet)
                for (int c = 0; c < 3; ++c) tm->wavelet[c].assign(t.wavelet[c], t.wavelet[c] + t.numWavelet);
        }
    }

private:
    const GailHipPlugin* api_;
    GailHipDecoder* d_;
};

class HipDevice : public Device {
public:
    HipDevice(const GailHipPlugin* api, GailHipDevice* d, const std::string& toolkit)
        : api_(api), d_(d), name_(std::string(api->device_name(d)) + " (HIP)"), tag_("hip") {
        // the COM tag: "hip" and the runtime's version, digits and dots, within the 9 bytes a
        // device tag may take ("HIP 7.2" -> "hip7.2")
        const size_t at = toolkit.find_first_of("0123456789");
        for (size_t i = at; i != std::string::npos && i < toolkit.size() && tag_.size() < 9; ++i) {
            const char ch = toolkit[i];
            if ((ch < '0' || ch > '9') && ch != '.') break;
            tag_ += ch;
        }
    }
    ~HipDevice() override { api_->device_destroy(d_); }
    std::string name() const override { return name_; }
    std::string codestreamTag() const override { return tag_; }
    std::unique_ptr<EncoderSlot> createSlot(const FrameLayout& L, const EncodeSettings& s) override {
        const PlainLayout pl(L);
        const GailHipEncodeSettings ps = plainSettings(s);
        GailHipEncoder* e = nullptr;
        check(api_, api_->encoder_create(d_, &pl.layout, &ps, &e));
        return std::make_unique<HipEncoderSlot>(api_, e, s.width, s.height);
    }
    uint64_t slotBytes(const FrameLayout& L, const EncodeSettings& s) const override {
        const PlainLayout pl(L);
        const GailHipEncodeSettings ps = plainSettings(s);
        return api_->encoder_bytes(d_, &pl.layout, &ps);
    }
    DeviceMemory memory() const override {
        GailHipMemory m{};
        DeviceMemory r;
        if (api_->device_memory(d_, &m) == GAIL_HIP_OK) {
            r.total = m.total;
            r.available = m.available;
            r.budgeted = m.budgeted != 0;
        }
        return r;
    }
    bool canDecode() const override { return true; }
    std::unique_ptr<DecoderSlot> createDecoderSlot(const FrameLayout& L, const DecodeSettings& s) override {
        const PlainLayout pl(L);
        const GailHipDecodeSettings ds{s.reduce, s.highBits ? 1 : 0};
        GailHipDecoder* d = nullptr;
        check(api_, api_->decoder_create(d_, &pl.layout, &ds, &d));
        return std::make_unique<HipDecoderSlot>(api_, d);
    }
    uint64_t decoderSlotBytes(const FrameLayout& L, const DecodeSettings& s) const override {
        const PlainLayout pl(L);
        const GailHipDecodeSettings ds{s.reduce, s.highBits ? 1 : 0};
        return api_->decoder_bytes(d_, &pl.layout, &ds);
    }

private:
    const GailHipPlugin* api_;
    GailHipDevice* d_;
    std::string name_, tag_;
};

}  // namespace

HipPluginInfo loadHipPlugin(const std::string& dir) {
    Plugin& p = plugin();
    std::call_once(p.once, [&] {
        p.info.path = dir + "/libgail_hip.so";
        void* h = dlopen(p.info.path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            p.info.problem = std::string("not loaded: ") + dlerror();
            return;
        }
        using Entry = const GailHipPlugin* (*)(uint32_t);
        const Entry entry = reinterpret_cast<Entry>(dlsym(h, GAIL_HIP_PLUGIN_ENTRY));
        const GailHipPlugin* api = entry ? entry(GAIL_HIP_PLUGIN_ABI) : nullptr;
        if (!entry) p.info.problem = "no " GAIL_HIP_PLUGIN_ENTRY " entry point";
        else if (!api) p.info.problem = "declines plugin ABI " + std::to_string(GAIL_HIP_PLUGIN_ABI);
        else if (api->abi != GAIL_HIP_PLUGIN_ABI || api->size != sizeof(GailHipPlugin) ||
                 api->layout != gail_hip_plugin_layout())
            p.info.problem = "built for another interface (ABI, table size or structure layout)";
        if (!p.info.problem.empty()) return;  // the library stays loaded; it is not used
        p.api = api;
        p.info.toolkit = api->toolkit;
    });
    return p.info;
}

std::vector<std::unique_ptr<Device>> createHipDevices(const std::string& dir, const std::vector<int>& select,
                                                      const Logger& log) {
    std::vector<std::unique_ptr<Device>> out;
    const HipPluginInfo info = loadHipPlugin(dir);
    const GailHipPlugin* api = plugin().api;
    if (!api) {
        if (log) log(LogLevel::Warning, "hip: plugin " + info.path + " " + info.problem);
        return out;
    }
    const int n = api->device_count();
    if (n < 0 && log) log(LogLevel::Warning, std::string("hip: ") + api->last_error());
    std::vector<int> want = select;
    if (want.empty() && n > 0) want.push_back(0);  // one GPU unless asked for more
    for (int i : want) {
        if (i < 0 || i >= n) continue;
        if (api->device_usable(i) != GAIL_HIP_OK) {
            if (log) log(LogLevel::Warning, "hip: device " + std::to_string(i) + " left out: " + api->last_error());
            continue;
        }
        GailHipDevice* d = nullptr;
        if (api->device_create(i, nullptr, nullptr, &d) != GAIL_HIP_OK) {
            if (log) log(LogLevel::Warning, "hip: device " + std::to_string(i) + " not set up: " + api->last_error());
            continue;
        }
        out.push_back(std::make_unique<HipDevice>(api, d, info.toolkit));
        if (log) log(LogLevel::Info, "hip: device " + std::to_string(i) + " " + out.back()->name());
    }
    return out;
}

}  // namespace gail
