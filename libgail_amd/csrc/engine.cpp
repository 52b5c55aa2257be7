// Device engine of the HIP backend (see engine.h). Host code compiled by hipcc.
#include "engine.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>

#include "codestream.h"
#include "errors.h"
#include "kernels_dec.h"

namespace gail_amd {

void hipCheck(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw DeviceOutOfMemory(std::string(what) + ": " + hipGetErrorString(e));
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

size_t InputFormat::bytes() const {
    switch (kind) {
    case RGB48: return (size_t)width * height * 6;
    case XyzPlanar: return (size_t)width * height * 6;
    default: {
        const size_t bps = bits > 8 ? 2 : 1;
        const size_t cw = (size_t)((aw + sx) >> sx), ch = (size_t)((ah + sy) >> sy);
        return ((size_t)aw * ah + 2 * cw * ch) * bps;
    }
    }
}

// ---------------------------------------------------------------- device context
DeviceContext& DeviceContext::get(int device) {
    static std::mutex m;
    static std::map<int, std::unique_ptr<DeviceContext>> ctx;
    std::lock_guard<std::mutex> l(m);
    auto& p = ctx[device];
    if (!p) p.reset(new DeviceContext(device));
    return *p;
}

DeviceContext::DeviceContext(int device) : dev_(device) {
    hipCheck(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    hipCheck(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    cus_ = prop.multiProcessorCount;
    name_ = prop.name;
    arch_ = prop.gcnArchName;
    hipCheck(t1_tables_init(), "Tier-1 tables");
    hipCheck(dec_tables_init(), "Tier-1 decoder tables");
    std::vector<float> t;
    buildLog2Table(t, kLog2Max);
    hipCheck(hipMalloc(&log2_, t.size() * sizeof(float)), "log2 table");
    hipCheck(hipMemcpy(log2_, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice), "log2 table upload");
    buildCodeThresholds(t);
    hipCheck(hipMalloc(&thr_, t.size() * sizeof(float)), "code thresholds");
    hipCheck(hipMemcpy(thr_, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice), "code thresholds upload");
    const OctaveSteps& o = octaveSteps();
    hipCheck(hipMalloc(&octM_, sizeof o.mantissa), "slope grid");
    hipCheck(hipMemcpy(octM_, o.mantissa, sizeof o.mantissa, hipMemcpyHostToDevice), "slope grid upload");
    hipCheck(hipMalloc(&octF_, sizeof o.first), "slope grid");
    hipCheck(hipMemcpy(octF_, o.first, sizeof o.first, hipMemcpyHostToDevice), "slope grid upload");
}

const DevRcTables& DeviceContext::rcTables(const Layout& L) {
    // keyed by the geometry and every subband's Mb (the only layout data the tables hold)
    std::string key = std::to_string(L.width) + "x" + std::to_string(L.height) + (L.fourk ? "k" : "s") +
                      std::to_string(L.levels) + ":";
    for (const Subband& b : L.bands) key += (char)('A' + b.Mb);
    std::lock_guard<std::mutex> lk(rcMutex_);
    auto it = rc_.find(key);
    if (it != rc_.end()) return it->second;
    const RcTables T = makeRcTables(L);
    DevRcTables d;
    RcPacket* pk = nullptr;
    RcPBand* pb = nullptr;
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipMalloc(&pk, std::max<size_t>(16, T.packets.size() * sizeof(RcPacket))), "packet table");
    hipCheck(hipMemcpy(pk, T.packets.data(), T.packets.size() * sizeof(RcPacket), hipMemcpyHostToDevice), "packet table");
    hipCheck(hipMalloc(&pb, std::max<size_t>(16, T.pbands.size() * sizeof(RcPBand))), "precinct table");
    hipCheck(hipMemcpy(pb, T.pbands.data(), T.pbands.size() * sizeof(RcPBand), hipMemcpyHostToDevice), "precinct table");
    d.packets = pk;
    d.pbands = pb;
    d.npackets = (uint32_t)T.packets.size();
    d.ntp = T.ntp;
    for (int k = 0; k < 6; ++k) d.tpComp[k] = T.tpComp[k];
    d.fixedBytes = T.fixedBytes;
    return rc_.emplace(key, d).first->second;
}

// ---------------------------------------------------------------- job
namespace {
constexpr size_t kAlign = 256;
size_t alignUp(size_t v, size_t a = kAlign) { return (v + a - 1) / a * a; }

struct Sizes {
    size_t pool = 0, out = 0;
};

Sizes t1Sizes(size_t coefs, uint32_t n, double perCoef) {
    Sizes z;
    z.pool = std::max<size_t>(64ull << 20, (size_t)(perCoef * (double)coefs));
    z.pool = std::max<size_t>(z.pool, (size_t)n * 1400);
    z.pool = alignUp(z.pool, 1 << 20);
    z.out = alignUp(z.pool / 2 + (size_t)n * 68);
    return z;
}
}  // namespace

template <class T> T* Job::devAlloc(size_t count) {
    void* p = nullptr;
    const size_t b = alignUp(std::max<size_t>(count * sizeof(T), 16));
    hipCheck(hipMalloc(&p, b), "device allocation");
    allocs_.push_back(p);
    bytes_ += b;
    return (T*)p;
}

namespace {
// Largest number of code-blocks and levels of any DCI layout (2K or 4K decomposition) of a
// frame no larger than W x H: block counts only grow with the frame.
struct SlotLimits {
    uint32_t blocks = 0, bands = 0;
    int levels = 0;
};
SlotLimits slotLimits(const SlotCapacity& c) {
    SlotLimits r;
    for (bool fourk : {false, true}) {
        const Layout L = makeGeometry(c.maxWidth, c.maxHeight, fourk, fourk ? 6 : 5, fourk ? 2 : 1);
        r.blocks = std::max<uint32_t>(r.blocks, (uint32_t)L.numCblks());
        r.bands = std::max<uint32_t>(r.bands, (uint32_t)L.bands.size());
        r.levels = std::max(r.levels, L.levels);
    }
    return r;
}
}  // namespace

bool SlotCapacity::fits(const FrameConfig& f) const {
    if (f.layout.width > maxWidth || f.layout.height > maxHeight) return false;
    if (f.input.kind != InputFormat::XyzPlanar && !colour) return false;
    return f.input.bytes() <= (size_t)maxWidth * maxHeight * 6;
}

size_t Job::estimateBytes(const std::vector<SlotCapacity>& slots, const JobOptions& opt) {
    size_t total = 0, coefs = 0;
    uint32_t n = 0, nb = 0;
    int lv = 0;
    for (const SlotCapacity& c : slots) {
        const size_t W = c.maxWidth, H = c.maxHeight, npix = W * H;
        total += alignUp(npix * 6) + (c.colour ? alignUp(npix * 6) + alignUp(65536 * 4) : 0);
        const CoefLayout cl = CoefLayout::of(W);
        total += alignUp(3 * cl.stride * H * 4) + alignUp(3 * cl.llStride * ((H + 1) / 2) * 4) +
                 alignUp(3 * ((W + 3) / 4) * ((H + 3) / 4) * 4);
        coefs += 3 * npix;
        const SlotLimits lim = slotLimits(c);
        n += lim.blocks;
        nb += lim.bands;
        lv = std::max(lv, lim.levels);
    }
    const Sizes z = t1Sizes(coefs, n, opt.poolBytesPerCoef);
    total += (size_t)n * (16 + 16 + 3 * kMaxPasses * 4 + kMaxPlanes * 8 + kMaxPasses * 16 + 4 + 8 + 8 + 4);
    total += (size_t)nb * sizeof(T1Band) + (size_t)lv * slots.size() * sizeof(DwtLevelDesc);
    // device rate control: hull points and their slope bins, chosen passes, slope ranges,
    // one histogram per slot
    total += (size_t)n * (kMaxPasses * (8 + 8 + 4 + 4) + 4 + 4) + (size_t)nb * 8;
    total += 16 * rcPartEntries(n, slots.size()) + 4 * rc_segsum_entries(slots.size());
    total += slots.size() * (6 * (size_t)(kRcMaxBins + 1) * 4 + sizeof(RcFrame));
    total += z.pool + z.out;
    return total;
}

static std::vector<SlotCapacity> exactSlots(const std::vector<FrameConfig>& frames) {
    std::vector<SlotCapacity> s;
    for (const FrameConfig& f : frames)
        s.push_back(SlotCapacity{f.layout.width, f.layout.height, f.input.kind != InputFormat::XyzPlanar});
    return s;
}

size_t Job::estimateBytes(const std::vector<FrameConfig>& frames, const JobOptions& opt) {
    return estimateBytes(exactSlots(frames), opt);
}

Job::Job(int device, std::vector<SlotCapacity> slots, const JobOptions& opt)
    : dev_(device), ctx_(DeviceContext::get(device)), opt_(opt), slots_(std::move(slots)) {
    if (slots_.empty()) throw std::logic_error("a job needs at least one frame slot");
    for (const SlotCapacity& c : slots_)
        if (c.maxWidth <= 0 || c.maxHeight <= 0 || c.maxWidth > 65535 || c.maxHeight > 65535)
            throw std::logic_error("invalid frame slot size");
    try {
        allocate();
    } catch (...) {
        release();
        throw;
    }
}

Job::Job(int device, std::vector<FrameConfig> frames, const JobOptions& opt) : Job(device, exactSlots(frames), opt) {
    bind(std::move(frames));
}

Job::~Job() { release(); }

void Job::release() {
    hipSetDevice(dev_);
    if (stream_) hipStreamSynchronize(stream_);
    if (upStream_) hipStreamSynchronize(upStream_);
    if (graph_) hipGraphExecDestroy(graph_);
    graph_ = nullptr;
    for (void* p : allocs_) hipFree(p);
    allocs_.clear();
    for (void* p : hostIn_)
        if (p) hipHostFree(p);
    hostIn_.clear();
    for (SlotDev& d : sd_) {
        if (d.hostLut) {
            hipHostFree(d.hostLut);
            d.hostLut = nullptr;
        }
        if (d.lut3) {
            hipFree(d.lut3);
            d.lut3 = nullptr;
            d.lut3Cap = 0;
        }
    }
    for (void** p : {(void**)&hFrames_, (void**)&hInfo_, (void**)&hPassLen_, (void**)&hPassDist_, (void**)&hData_,
                     (void**)&hSmall_, (void**)&hBlocks_, (void**)&hBands_, (void**)&hLevels_, (void**)&hBandScale_,
                     (void**)&hRc_})
        if (*p) {
            hipHostFree(*p);
            *p = nullptr;
        }
    for (auto& e : ev_)
        if (e) {
            hipEventDestroy(e);
            e = nullptr;
        }
    if (evDone_) hipEventDestroy(evDone_);
    evDone_ = nullptr;
    if (evUp_) hipEventDestroy(evUp_);
    evUp_ = nullptr;
    if (stream_) hipStreamDestroy(stream_);
    stream_ = nullptr;
    if (upStream_) hipStreamDestroy(upStream_);
    upStream_ = nullptr;
    upPending_ = false;
    bytes_ = 0;
}

static void hostAlloc(void** p, size_t bytes) {
    hipCheck(hipHostMalloc(p, std::max<size_t>(bytes, 16), hipHostMallocDefault), "pinned host memory");
}

void Job::allocate() {
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
    hipCheck(hipStreamCreateWithFlags(&upStream_, hipStreamNonBlocking), "stream");
    for (auto& e : ev_) hipCheck(hipEventCreate(&e), "event");
    hipCheck(hipEventCreateWithFlags(&evDone_, hipEventDisableTiming), "event");
    hipCheck(hipEventCreateWithFlags(&evUp_, hipEventDisableTiming), "event");
    const int ns = (int)slots_.size();
    sd_.resize(ns);
    hostIn_.assign(ns, nullptr);
    slope_.assign(ns, 0.f);
    for (int i = 0; i < ns; ++i) {
        const SlotCapacity& c = slots_[i];
        SlotDev& d = sd_[i];
        const size_t W = c.maxWidth, H = c.maxHeight;
        d.pixCap = W * H;
        d.inCap = 6 * d.pixCap;
        d.in = devAlloc<uint8_t>(d.inCap);
        if (c.colour) {
            d.codesBuf = devAlloc<uint16_t>(3 * d.pixCap);
            d.lut = devAlloc<float>(65536);
            hostAlloc((void**)&d.hostLut, 65536 * sizeof(float));
        }
        // the padded layout of the widest frame is the widest of every frame that fits
        const CoefLayout cl = CoefLayout::of(W);
        d.coef = devAlloc<float>(3 * cl.stride * H);
        d.llA = devAlloc<float>(3 * cl.llStride * ((H + 1) / 2));
        d.llB = devAlloc<float>(3 * ((W + 3) / 4) * ((H + 3) / 4));
        const SlotLimits lim = slotLimits(c);
        blockCap_ += lim.blocks;
        bandCap_ += lim.bands;
        levelCap_ = std::max(levelCap_, lim.levels);
        coefCap_ += 3 * d.pixCap;
    }
    if (bandCap_ > 65535 || ns > 65535) throw std::logic_error("job too large");
    dBlocks_ = devAlloc<T1Block>(blockCap_);
    hostAlloc((void**)&hBlocks_, sizeof(T1Block) * blockCap_);
    dBands_ = devAlloc<T1Band>(bandCap_);
    hostAlloc((void**)&hBands_, sizeof(T1Band) * bandCap_);
    dFrames_ = devAlloc<T1FrameRef>(ns);
    hostAlloc((void**)&hFrames_, sizeof(T1FrameRef) * ns);
    dLevels_ = devAlloc<DwtLevelDesc>((size_t)levelCap_ * ns);
    hostAlloc((void**)&hLevels_, sizeof(DwtLevelDesc) * levelCap_ * ns);
    dBandScale_ = devAlloc<double>(bandCap_);
    hostAlloc((void**)&hBandScale_, sizeof(double) * bandCap_);
    dRc_ = devAlloc<RcFrame>(ns);
    hostAlloc((void**)&hRc_, sizeof(RcFrame) * ns);
    std::memset(hRc_, 0, sizeof(RcFrame) * ns);
    dHullSlope_ = devAlloc<double>((size_t)blockCap_ * kMaxPasses);
    dHullGain_ = devAlloc<double>((size_t)blockCap_ * kMaxPasses);
    dHullSel_ = devAlloc<uint32_t>((size_t)blockCap_ * kMaxPasses);
    dHullKbin_ = devAlloc<int32_t>((size_t)blockCap_ * kMaxPasses);
    dHullN_ = devAlloc<uint32_t>(blockCap_);
    dRcSel_ = devAlloc<uint32_t>(blockCap_);
    dRcPart_ = devAlloc<unsigned long long>(2 * rcPartEntries(blockCap_, ns));
    dRcSegsum_ = devAlloc<uint32_t>(rc_segsum_entries(ns));
    // the rate control leaves its histogram zero after every run; it starts zero
    dRcHist_ = devAlloc<uint32_t>((size_t)ns * 6 * (kRcMaxBins + 1));
    hipCheck(hipMemset(dRcHist_, 0, sizeof(uint32_t) * ns * 6 * (kRcMaxBins + 1)), "histogram clear");

    const uint32_t n = blockCap_;
    dInfo_ = devAlloc<T1Info>(n);
    dPassEnd_ = devAlloc<uint32_t>((size_t)n * kMaxPasses);
    dPassDist_ = devAlloc<int32_t>((size_t)n * kMaxPasses);
    dPassLen_ = devAlloc<uint32_t>((size_t)n * kMaxPasses);
    dChunks_ = devAlloc<uint2>((size_t)n * kMaxPlanes);
    dSnaps_ = devAlloc<MqSnapshot>((size_t)n * kMaxPasses);
    dPerm_ = devAlloc<uint32_t>(n);
    dHist_ = devAlloc<uint32_t>(128);
    dOutOff_ = devAlloc<unsigned long long>(n);
    dDataOff_ = devAlloc<unsigned long long>(n);
    dPartial_ = devAlloc<unsigned long long>(n / 2048 + 4);
    dCodeLen_ = devAlloc<uint32_t>(n);
    dCounters_ = devAlloc<JobCounters>(1);
    dWork_ = devAlloc<uint32_t>(4);
    hostAlloc((void**)&hSmall_, 256);
    hostAlloc((void**)&hInfo_, (size_t)n * 20);
    // result buffers start at typical sizes and grow when a run needs more
    hPassCap_ = (size_t)n * 32;
    hostAlloc((void**)&hPassLen_, hPassCap_ * 4);
    hostAlloc((void**)&hPassDist_, hPassCap_ * 4);
    hDataCap_ = std::max<size_t>(16ull << 20, coefCap_ / 2);
    hostAlloc((void**)&hData_, hDataCap_);
    const Sizes z = t1Sizes(coefCap_, n, opt_.poolBytesPerCoef);
    poolCap_ = z.pool;
    outCap_ = z.out;
    dPool_ = devAlloc<uint8_t>(poolCap_);
    dOut_ = devAlloc<uint8_t>(outCap_);
}

void Job::bind(std::vector<FrameConfig> frames) {
    const int nf = (int)frames.size();
    if (nf < 1 || nf > numSlots()) throw std::logic_error("more frames than the job has slots");
    for (int f = 0; f < nf; ++f) {
        const FrameConfig& fc = frames[f];
        if (fc.input.width != fc.layout.width || fc.input.height != fc.layout.height)
            throw std::logic_error("input format does not match the frame size");
        if (!slots_[f].fits(fc)) throw std::logic_error("frame does not fit its slot");
    }
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipStreamSynchronize(stream_), "previous run");  // the staging tables are reused
    // Until the new tables are complete nothing is bound (a failure leaves the job unbound).
    frames_.clear();
    nblocks_ = 0;
    launched_ = false;
    uint32_t nblocks = 0, nbands = 0;
    int maxLevels = 0;
    bool anyRc = false;
    uint32_t rcMaxBlocks = 0;
    for (int f = 0; f < nf; ++f) {
        const FrameConfig& fc = frames[f];
        const Layout& L = fc.layout;
        SlotDev& d = sd_[f];
        const size_t W = L.width, H = L.height;
        d.npix = W * H;
        const CoefLayout cl = CoefLayout::of(W);
        d.coefStride = cl.stride;
        d.coefPlane = cl.stride * H;
        d.hiOff = cl.hiOff;
        d.llStride = cl.llStride;
        d.inBytes = fc.input.bytes();
        d.codes = fc.input.kind == InputFormat::XyzPlanar ? (uint16_t*)d.in : d.codesBuf;
        if (fc.input.kind != InputFormat::XyzPlanar) {
            const ColorSpec& cs = fc.color;
            if (!d.lutValid || !cs.sameTables(d.lutSpec)) {
                std::vector<float> lut;
                buildColorTables(cs, lut, d.params);
                if (cs.lut3d && cs.lut3dSize > 1) {
                    // an app 3D LUT lives beside the slot, grown to the largest one seen
                    const size_t n = cs.lut3d->size();
                    if (n > d.lut3Cap) {
                        if (d.lut3) hipFree(d.lut3);
                        d.lut3 = nullptr;
                        d.lut3Cap = 0;
                        hipCheck(hipMalloc((void**)&d.lut3, n * sizeof(float)), "3D LUT");
                        d.lut3Cap = n;
                    }
                    hipCheck(hipMemcpyAsync(d.lut3, cs.lut3d->data(), n * sizeof(float), hipMemcpyHostToDevice, stream_),
                             "3D LUT upload");
                }
                std::memcpy(d.hostLut, lut.data(), 65536 * sizeof(float));
                hipCheck(hipMemcpyAsync(d.lut, d.hostLut, 65536 * sizeof(float), hipMemcpyHostToDevice, stream_),
                         "EOTF table upload");
                d.lutSpec = cs;
                d.lutValid = true;
            }
        }
        d.firstBlock = nblocks;
        d.numBlocks = (uint32_t)L.numCblks();
        if (nblocks + d.numBlocks > blockCap_ || nbands + L.bands.size() > bandCap_)
            throw std::logic_error("frame layout exceeds the job's block capacity");
        const uint32_t bandBase = nbands;
        for (const Subband& b : L.bands) hBandScale_[nbands + (&b - L.bands.data())] = (double)b.delta * b.delta * b.weight;
        for (const Subband& b : L.bands)
            hBands_[nbands++] = T1Band{1.0f / b.delta, ((1u << b.Mb) << 6) - 1u, (float)((double)b.delta * b.delta * b.weight),
                                       (uint32_t)b.comp | ((uint32_t)b.Mb << 8)};
        T1Block* out = hBlocks_ + nblocks;
        for (int i = 0; i < L.numCblks(); ++i) {
            const CodeBlock& cb = L.cblks[i];
            const Subband& b = L.bands[cb.band];
            T1Block t{};
            // level 0's HL and HH are the only bands right of column ceil(W / 2)
            size_t x = (size_t)(b.px0 + cb.x0);
            if (x >= (W + 1) / 2) x += d.hiOff - (W + 1) / 2;
            t.coef_off = (uint32_t)((size_t)b.comp * d.coefPlane + (size_t)(b.py0 + cb.y0) * d.coefStride + x);
            t.frame = (uint16_t)f;
            t.band = (uint16_t)(bandBase + cb.band);
            t.w = (uint8_t)cb.w;
            t.h = (uint8_t)cb.h;
            t.orient = (uint8_t)b.orient;
            t.frame_block = (uint32_t)i;
            out[i] = t;
        }
        RcFrame& rc = hRc_[f];
        std::memset(&rc, 0, sizeof rc);
        if (fc.rcMaxTotal) {
            const DevRcTables& T = ctx_.rcTables(L);
            rc.packets = T.packets;
            rc.pbands = T.pbands;
            rc.npackets = T.npackets;
            rc.ntp = T.ntp;
            rc.first_block = nblocks;
            rc.nblocks = d.numBlocks;
            rc.fixed_bytes = T.fixedBytes;
            rc.max_total = fc.rcMaxTotal;
            rc.max_comp = fc.rcMaxComp ? fc.rcMaxComp : fc.rcMaxTotal;
            for (int k = 0; k < 6; ++k) rc.tp_comp[k] = T.tpComp[k];
            anyRc = true;
            rcMaxBlocks = std::max(rcMaxBlocks, d.numBlocks);
        }
        nblocks += d.numBlocks;
        maxLevels = std::max(maxLevels, L.levels);
        hFrames_[f] = T1FrameRef{d.coef, (uint32_t)d.coefStride, 0.f};
    }
    if (maxLevels > levelCap_) throw std::logic_error("frame has more levels than its slot");
    hipStream_t s = stream_;
    hipCheck(hipMemcpyAsync(dBlocks_, hBlocks_, sizeof(T1Block) * nblocks, hipMemcpyHostToDevice, s), "block table");
    hipCheck(hipMemcpyAsync(dBands_, hBands_, sizeof(T1Band) * nbands, hipMemcpyHostToDevice, s), "band table");
    hipCheck(hipMemcpyAsync(dBandScale_, hBandScale_, sizeof(double) * nbands, hipMemcpyHostToDevice, s), "band scales");
    hipCheck(hipMemcpyAsync(dRc_, hRc_, sizeof(RcFrame) * nf, hipMemcpyHostToDevice, s), "rate-control frames");
    anyRc_ = anyRc;
    rcMaxBlocks_ = rcMaxBlocks;
    nblocks_ = nblocks;
    numLevels_ = maxLevels;
    frames_ = std::move(frames);
    planLevels();
    frameDataBase_.assign(nf, 0);
    framePassBase_.assign(nf, 0);
    launched_ = false;
    if (graph_) {  // the captured chain belongs to the previous composition
        hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
}

// Replaces the symbol pool and / or the codeword buffer with ones twice as large (after a
// run overflowed them); the chain is then run again from context modelling.
void Job::growT1(bool pool, bool out) {
    hipCheck(hipStreamSynchronize(stream_), "sync");
    auto replace = [&](uint8_t*& p, size_t& cap) {
        auto it = std::find(allocs_.begin(), allocs_.end(), (void*)p);
        if (it != allocs_.end()) allocs_.erase(it);
        hipFree(p);
        bytes_ -= alignUp(cap);
        cap *= 2;
        p = devAlloc<uint8_t>(cap);
    };
    // Both grow together: the dense results live in the pool behind room for every
    // codeword, so the pool must stay larger than the codeword buffer.
    (void)pool;
    (void)out;
    replace(dPool_, poolCap_);
    replace(dOut_, outCap_);
    if (graph_) {
        hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
}

// DWT level descriptors of the bound frames for the current tile shape: level l of every
// frame that has it, tiles numbered across frames; queued to the device on the stream.
void Job::planLevels() {
    const int nf = numFrames();
    levelCount_.assign(numLevels_, 0);
    levelTiles_.assign(numLevels_, 0);
    // each level's shape: the job's, but the strips need LL regions of 2 x 2 samples at least
    levelShape_.assign(numLevels_, opt_.dwtTile);
    if (opt_.dwtTile >= 7)  // strips with rows loaded ahead at level 0 only, LDS tiles below it
        for (int l = 1; l < numLevels_; ++l) levelShape_[l] = 3;
    if (opt_.dwtTile >= 4)
        for (int f = 0; f < nf; ++f) {
            const Layout& L = frames_[f].layout;
            int w = L.width, h = L.height;
            // the buffer-addressed strips mirror rows by one reflection (frames of 11 rows or
            // more) and address a frame's three planes with 32-bit byte offsets
            if (levelShape_[0] >= 9 &&
                (h < 11 || (uint64_t)3 * (uint64_t)sd_[f].coefPlane * sizeof(float) >= (uint64_t)INT32_MAX))
                levelShape_[0] = 8;
            for (int l = 0; l < L.levels; ++l) {
                if (w < 2 || h < 2) levelShape_[l] = 3;
                w = (w + 1) / 2;
                h = (h + 1) / 2;
            }
        }
    for (int f = 0; f < nf; ++f) {
        const Layout& L = frames_[f].layout;
        SlotDev& d = sd_[f];
        int w = L.width, h = L.height;
        for (int l = 0; l < L.levels; ++l) {
            DwtLevelDesc ds{};
            const int w2 = (w + 1) / 2, h2 = (h + 1) / 2;
            // level 0's LL goes to llA in rows of d.llStride; the deeper levels' LL planes
            // are packed (row stride = width)
            if (l == 0) {
                ds.src = d.codes;
                ds.src_stride = (uint32_t)L.width;
                ds.src_plane = (uint32_t)d.npix;
            } else {
                const uint32_t st = l == 1 ? (uint32_t)d.llStride : (uint32_t)w;
                ds.src = ((l - 1) % 2 == 0) ? d.llA : d.llB;
                ds.src_stride = st;
                ds.src_plane = st * (uint32_t)h;
            }
            ds.coef = d.coef;
            ds.coef_stride = (uint32_t)d.coefStride;
            ds.coef_plane = (uint32_t)d.coefPlane;
            ds.hi_off = l == 0 ? (uint32_t)d.hiOff : (uint32_t)w2;
            if (l == L.levels - 1) {
                ds.ll_dst = d.coef;
                ds.ll_stride = (uint32_t)d.coefStride;
                ds.ll_plane = (uint32_t)d.coefPlane;
            } else {
                const uint32_t st = l == 0 ? (uint32_t)d.llStride : (uint32_t)w2;
                ds.ll_dst = (l % 2 == 0) ? d.llA : d.llB;
                ds.ll_stride = st;
                ds.ll_plane = st * (uint32_t)h2;
            }
            ds.w = w;
            ds.h = h;
            int tx, ty;
            dwt_tile_size(levelShape_[l], tx, ty);
            ds.tiles_x = (w + tx - 1) / tx;
            ds.tile_base = (uint32_t)levelTiles_[l];
            levelTiles_[l] += ds.tiles_x * ((h + ty - 1) / ty);
            hLevels_[(size_t)l * numSlots() + levelCount_[l]++] = ds;
            w = w2;
            h = h2;
        }
    }
    hipCheck(hipMemcpyAsync(dLevels_, hLevels_, sizeof(DwtLevelDesc) * numLevels_ * numSlots(), hipMemcpyHostToDevice,
                            stream_),
             "DWT descriptors");
    if (graph_) {  // the captured chain has the old launch shapes
        hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
}

void Job::enqueueDwt(hipStream_t s) {
    for (int l = 0; l < numLevels_; ++l) {
        hipCheck(launch_dwt_level(l == 0, levelShape_[l], dLevels_ + (size_t)l * numSlots(), levelCount_[l], levelTiles_[l], s),
                 "DWT level");
        if (l == 0 && !capturing_) hipCheck(hipEventRecord(ev_[8], s), "event");
    }
}

JobOptions Job::autotune(int reps, TuneTimes* out) {
    if (frames_.empty()) throw std::logic_error("no frames bound to the job");
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipStreamSynchronize(stream_), "previous run");
    hipCheck(hipStreamSynchronize(upStream_), "frame uploads");
    upPending_ = false;
    launched_ = false;
    reps = std::max(1, reps);
    hipStream_t s = stream_;
    for (int f = 0; f < numFrames(); ++f) hFrames_[f].slope_thresh = slope_[f];
    hipCheck(hipMemcpyAsync(dFrames_, hFrames_, sizeof(T1FrameRef) * numFrames(), hipMemcpyHostToDevice, s), "frame table");
    auto elapsed = [&](hipEvent_t a, hipEvent_t b) {
        hipCheck(hipEventSynchronize(b), "autotune");
        float ms = 0;
        hipCheck(hipEventElapsedTime(&ms, a, b), "autotune");
        return (double)ms;
    };
    TuneTimes t;
    // DWT tile shape: the whole DWT of the bound frames, best of reps
    int bestTile = opt_.dwtTile;
    double bestDwt = 1e30;
    for (int shape = 0; shape < kDwtShapes; ++shape) {
        opt_.dwtTile = shape;
        planLevels();
        double best = 1e30;
        for (int r = 0; r <= reps; ++r) {  // the first run warms up
            hipCheck(hipEventRecord(ev_[2], s), "event");
            enqueueDwt(s);
            hipCheck(hipEventRecord(ev_[3], s), "event");
            const double ms = elapsed(ev_[2], ev_[3]);
            if (r) best = std::min(best, ms);
        }
        t.dwt[shape] = best;
        if (best < bestDwt) {
            bestDwt = best;
            bestTile = shape;
        }
    }
    opt_.dwtTile = bestTile;
    planLevels();
    enqueueDwt(s);  // the planes the Tier-1 runs below read
    // context modelling waves per workgroup, then MQ workgroup size (whole Tier-1 runs)
    const int waves[4] = {1, 2, 4, 8}, groups[3] = {64, 128, 256};
    int bestWaves = opt_.cmWaves, bestGroup = opt_.mqGroup, bestRead = opt_.mqRead;
    double bestCm = 1e30, bestMq = 1e30;
    auto runT1 = [&]() {  // a run that overflowed its buffers skipped work: grow and repeat
        for (int attempt = 0;; ++attempt) {
            hipCheck(hipEventRecord(ev_[3], s), "event");
            enqueueT1(s);
            hipCheck(hipStreamSynchronize(s), "autotune");
            const uint32_t flags = ((const JobCounters*)hSmall_)->flags;
            if (flags & 4u) throw std::runtime_error("code-block with more than 24 magnitude bit-planes");
            if (!(flags & 3u)) return;
            if (attempt >= 6) throw std::runtime_error("Tier-1 buffers keep overflowing");
            growT1(true, true);
        }
    };
    for (int i = 0; i < 4; ++i) {
        opt_.cmWaves = waves[i];
        double best = 1e30;
        for (int r = 0; r <= reps; ++r) {
            runT1();
            const double ms = elapsed(ev_[3], ev_[4]);
            if (r) best = std::min(best, ms);
        }
        t.cm[i] = best;
        if (best < bestCm) {
            bestCm = best;
            bestWaves = waves[i];
        }
    }
    opt_.cmWaves = bestWaves;
    for (int i = 0; i < 6; ++i) {
        opt_.mqGroup = groups[i % 3];
        opt_.mqRead = i < 3 ? 16 : 64;
        double best = 1e30;
        for (int r = 0; r <= reps; ++r) {
            runT1();
            const double ms = elapsed(ev_[4], ev_[5]);
            if (r) best = std::min(best, ms);
        }
        t.mq[i] = best;
        if (best < bestMq) {
            bestMq = best;
            bestGroup = opt_.mqGroup;
            bestRead = opt_.mqRead;
        }
    }
    opt_.mqGroup = bestGroup;
    opt_.mqRead = bestRead;
    if (graph_) {
        hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    if (out) *out = t;
    return opt_;
}

void* Job::hostInput(int i) {
    if (i < 0 || i >= numSlots()) throw std::logic_error("no such frame slot");
    if (!hostIn_[i]) hostAlloc(&hostIn_[i], sd_[i].inCap);
    return hostIn_[i];
}

void Job::upload(int i, const void* frame) {
    if (i < 0 || i >= numFrames()) throw std::logic_error("no frame bound to this slot");
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipStreamSynchronize(stream_), "previous run");
    hipCheck(hipMemcpy(sd_[i].in, frame, sd_[i].inBytes, hipMemcpyHostToDevice), "frame upload");
}

void Job::uploadFrom(int i, const void* src) {
    if (i < 0 || i >= numFrames()) throw std::logic_error("no frame bound to this slot");
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    if (launched_) hipCheck(hipStreamSynchronize(stream_), "previous run");
    // host or device source: the runtime tells them apart (a device source stays in HBM)
    const bool second = (i & 1) != 0;
    hipCheck(hipMemcpyAsync(sd_[i].in, src, sd_[i].inBytes, hipMemcpyDefault, second ? upStream_ : stream_),
             "frame upload");
    if (second) upPending_ = true;
}

void Job::inputsReady() {
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipStreamSynchronize(stream_), "frame uploads");
    hipCheck(hipStreamSynchronize(upStream_), "frame uploads");
}

void Job::uploadAsync(int i) {
    if (i < 0 || i >= numFrames()) throw std::logic_error("no frame bound to this slot");
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    hipCheck(hipMemcpyAsync(sd_[i].in, hostInput(i), sd_[i].inBytes, hipMemcpyHostToDevice, stream_), "frame upload");
}

// Launches of at most this many blocks code their longest blocks one per wave: with a
// batch's worth of blocks the lane waves fill the machine and the scalar coder only competes.
constexpr uint32_t kMqUniformMaxBlocks = 65536;

void Job::enqueueT1(hipStream_t s) {
    const uint32_t n = nblocks_;
    hipCheck(launch_clear_counters(dCounters_, dWork_, 4, s), "clear counters");
    CmParams cm{};
    cm.blocks = dBlocks_;
    cm.bands = dBands_;
    cm.frames = dFrames_;
    cm.nblocks = n;
    cm.max_passes = kMaxPasses;
    cm.info = dInfo_;
    cm.pass_end = dPassEnd_;
    cm.pass_dist = dPassDist_;
    cm.chunks = dChunks_;
    cm.pool = dPool_;
    cm.pool_cap = poolCap_;
    cm.counters = dCounters_;
    cm.work_next = dWork_;
    cm.log2tab = ctx_.log2Table();
    const int waves = opt_.cmWaves;
    const uint32_t pairs = (n + 1) / 2;
    const int cmGroups = (int)std::max<uint32_t>(
        1, std::min<uint32_t>((pairs + waves - 1) / waves, (uint32_t)(ctx_.computeUnits() * opt_.cmGroupsPerCU)));
    hipCheck(launch_t1_cm(cm, waves, cmGroups, s), "context modelling");
    if (!capturing_) hipCheck(hipEventRecord(ev_[4], s), "event");
    hipCheck(launch_sort_blocks(dInfo_, n, dHist_, dPerm_, dCounters_, s), "block ordering");
    hipCheck(launch_out_offsets(dInfo_, n, dOutOff_, dPartial_, s), "codeword offsets");
    MqParams mq{};
    mq.info = dInfo_;
    mq.pass_end = dPassEnd_;
    mq.chunks = dChunks_;
    mq.pool = dPool_;
    mq.perm = dPerm_;
    mq.out_off = dOutOff_;
    mq.out = dOut_;
    mq.out_cap = outCap_;
    mq.snaps = dSnaps_;
    mq.pass_len = dPassLen_;
    mq.code_len = dCodeLen_;
    mq.counters = dCounters_;
    mq.nblocks = n;
    mq.work_next = dWork_ + 1;
    mq.nlong = n <= kMqUniformMaxBlocks ? std::min<uint32_t>(n, (uint32_t)(ctx_.computeUnits() * opt_.mqLongPerCU)) : 0u;
    // enough waves for one long block each, where the grid cap allows
    const uint32_t mqWaves = (uint32_t)opt_.mqGroup / 64;
    const uint32_t mqWant = std::max<uint32_t>((n + opt_.mqGroup - 1) / opt_.mqGroup, (mq.nlong + mqWaves - 1) / mqWaves);
    const int mqGroups = (int)std::max<uint32_t>(
        1, std::min<uint32_t>(mqWant, (uint32_t)(ctx_.computeUnits() * opt_.mqGroupsPerCU)));
    hipCheck(launch_t1_mq(mq, opt_.mqGroup, opt_.mqRead, mqGroups, s), "MQ coding");
    if (!capturing_) hipCheck(hipEventRecord(ev_[5], s), "event");
    if (anyRc_) {
        RcParams rp{};
        rp.frames = dRc_;
        rp.blocks = dBlocks_;
        rp.bands = dBands_;
        rp.band_scale = dBandScale_;
        rp.info = dInfo_;
        rp.pass_len = dPassLen_;
        rp.pass_dist = dPassDist_;
        rp.code_len = dCodeLen_;
        rp.stride = blockCap_;
        rp.h_slope = dHullSlope_;
        rp.h_gain = dHullGain_;
        rp.h_sel = dHullSel_;
        rp.h_kbin = dHullKbin_;
        rp.hull_n = dHullN_;
        rp.part = dRcPart_;
        rp.hist = dRcHist_;
        rp.segsum = dRcSegsum_;
        rp.mantissa = ctx_.octaveMantissa();
        rp.first = ctx_.octaveFirst();
        rp.sel = dRcSel_;
        rp.counters = dCounters_;
        hipCheck(launch_rc(rp, numFrames(), rcMaxBlocks_, s), "rate control");
    }
    hipCheck(launch_compact_offsets(dInfo_, dCodeLen_, n, dDataOff_, nullptr, dPartial_, s), "compaction offsets");
    // dense results live in the (now consumed) symbol pool
    const size_t dataBytes = alignUp(outCap_);
    dDenseData_ = dPool_;
    dDenseInfo_ = (uint32_t*)(dPool_ + dataBytes);
    dDensePassLen_ = (uint32_t*)(dPool_ + dataBytes + alignUp((size_t)n * 20));
    dDensePassDist_ = (int32_t*)(dPool_ + dataBytes + alignUp((size_t)n * 20) + alignUp((size_t)n * kMaxPasses * 4));
    CompactParams cp{};
    cp.info = dInfo_;
    cp.code_len = dCodeLen_;
    cp.out_off = dOutOff_;
    cp.out = dOut_;
    cp.pass_len = dPassLen_;
    cp.pass_dist = dPassDist_;
    cp.data_off = dDataOff_;
    cp.dense_info = dDenseInfo_;
    cp.dense_pass_len = dDensePassLen_;
    cp.dense_pass_dist = dDensePassDist_;
    cp.dense_data = dDenseData_;
    cp.counters = dCounters_;
    cp.nblocks = n;
    hipCheck(launch_compact(cp, s), "compaction");
    if (!capturing_) hipCheck(hipEventRecord(ev_[6], s), "event");
    // flags and totals for finish()
    const int groups = (int)((n + 2047) / 2048);
    hipCheck(hipMemcpyAsync(hSmall_, dCounters_, sizeof(JobCounters), hipMemcpyDeviceToHost, s), "counters");
    hipCheck(hipMemcpyAsync(hSmall_ + 16, dPartial_ + groups, 8, hipMemcpyDeviceToHost, s), "totals");
}

void Job::enqueueChain(bool uploadInputs) {
    hipStream_t s = stream_;
    if (!capturing_) hipCheck(hipEventRecord(ev_[0], s), "event");
    if (uploadInputs)
        for (int f = 0; f < numFrames(); ++f)
            hipCheck(hipMemcpyAsync(sd_[f].in, hostInput(f), sd_[f].inBytes, hipMemcpyHostToDevice, s), "frame upload");
    hipCheck(hipMemcpyAsync(dFrames_, hFrames_, sizeof(T1FrameRef) * numFrames(), hipMemcpyHostToDevice, s), "frame table");
    if (!capturing_) hipCheck(hipEventRecord(ev_[1], s), "event");
    for (int f = 0; f < numFrames(); ++f) {
        const InputFormat& in = frames_[f].input;
        if (in.kind == InputFormat::XyzPlanar) continue;
        ColorDesc c{};
        c.src = sd_[f].in;
        c.codes = sd_[f].codes;
        c.lut = sd_[f].lut;
        c.thresholds = ctx_.thresholds();
        std::memcpy(c.params, sd_[f].params, sizeof c.params);
        c.lut3 = sd_[f].lut3;
        c.lut3n = (int32_t)sd_[f].params[38];
        c.width = in.width;
        c.height = in.height;
        c.kind = in.kind == InputFormat::RGB48 ? 0 : 1;
        c.aw = in.aw;
        c.ah = in.ah;
        c.ox = in.ox;
        c.oy = in.oy;
        c.sx = in.sx;
        c.sy = in.sy;
        c.bits = in.bits;
        c.rgb = in.rgb;
        c.full_range = in.fullRange;
        c.chroma = in.chroma;
        const double kg = 1.0 - in.kr - in.kb;
        c.crR = (float)(2 * (1 - in.kr));
        c.cbB = (float)(2 * (1 - in.kb));
        c.cbG = (float)(2 * in.kb * (1 - in.kb) / kg);
        c.crG = (float)(2 * in.kr * (1 - in.kr) / kg);
        hipCheck(launch_color(c, s), "colour");
    }
    if (!capturing_) hipCheck(hipEventRecord(ev_[2], s), "event");
    enqueueDwt(s);
    if (!capturing_) hipCheck(hipEventRecord(ev_[3], s), "event");
    enqueueT1(s);
}

void Job::launch(bool uploadInputs) {
    if (frames_.empty()) throw std::logic_error("no frames bound to the job");
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    // hFrames_ is staged from the host: a run still in flight may be reading it
    if (launched_) hipCheck(hipStreamSynchronize(stream_), "previous run");
    if (upPending_) {  // the run reads what uploadFrom queued on the second stream
        hipCheck(hipEventRecord(evUp_, upStream_), "event");
        hipCheck(hipStreamWaitEvent(stream_, evUp_, 0), "stream wait");
        upPending_ = false;
    }
    for (int f = 0; f < numFrames(); ++f) hFrames_[f].slope_thresh = slope_[f];
    launched_ = true;
    if (uploadInputs)  // staging buffers exist before any capture records copies from them
        for (int f = 0; f < numFrames(); ++f) hostInput(f);
    if (opt_.useGraph) {
        if (!graph_ || graphUpload_ != uploadInputs) {
            if (graph_) hipGraphExecDestroy(graph_);
            graph_ = nullptr;
            hipGraph_t g = nullptr;
            hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "graph capture");
            capturing_ = true;
            try {
                enqueueChain(uploadInputs);
                capturing_ = false;
            } catch (...) {
                capturing_ = false;
                hipStreamEndCapture(stream_, &g);
                if (g) hipGraphDestroy(g);
                throw;
            }
            hipCheck(hipStreamEndCapture(stream_, &g), "graph capture");
            hipCheck(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0), "graph instantiate");
            hipGraphDestroy(g);
            graphUpload_ = uploadInputs;
        }
        hipCheck(hipEventRecord(ev_[0], stream_), "event");
        hipCheck(hipGraphLaunch(graph_, stream_), "graph launch");
        hipCheck(hipEventRecord(ev_[6], stream_), "event");
        hipCheck(hipEventRecord(evDone_, stream_), "event");
        return;
    }
    enqueueChain(uploadInputs);
    hipCheck(hipEventRecord(evDone_, stream_), "event");
}

bool Job::done() const {
    if (!launched_) return true;
    const hipError_t e = hipEventQuery(evDone_);
    if (e == hipErrorNotReady) return false;
    hipCheck(e, "encode");
    return true;
}

void Job::finish(StageTimes* times) {
    if (!launched_) throw std::logic_error("finish() without a launched run");
    hipCheck(hipStreamSynchronize(stream_), "encode");
    launched_ = false;
    int reruns = 0;
    for (;;) {
        const JobCounters* c = (const JobCounters*)hSmall_;
        const uint32_t flags = c->flags;
        if (flags & 4u) throw std::runtime_error("code-block with more than 24 magnitude bit-planes");
        if (!(flags & 3u)) break;
        if (++reruns > 6) throw std::runtime_error("Tier-1 buffers keep overflowing");
        growT1((flags & 1u) != 0, (flags & 2u) != 0);
        enqueueT1(stream_);
        hipCheck(hipStreamSynchronize(stream_), "encode (re-run)");
    }
    const unsigned long long packed = hSmall_[16];
    numDataTotal_ = (size_t)(packed & ((1ull << 36) - 1));
    numPassesTotal_ = (size_t)(packed >> 36);
    if (numPassesTotal_ > hPassCap_) {
        hipHostFree(hPassLen_);
        hipHostFree(hPassDist_);
        hPassLen_ = nullptr;
        hPassDist_ = nullptr;
        hPassCap_ = std::min<size_t>((size_t)blockCap_ * kMaxPasses, numPassesTotal_ + (numPassesTotal_ >> 2));
        hostAlloc((void**)&hPassLen_, hPassCap_ * 4);
        hostAlloc((void**)&hPassDist_, hPassCap_ * 4);
    }
    if (numDataTotal_ > hDataCap_) {
        hipHostFree(hData_);
        hData_ = nullptr;
        hDataCap_ = numDataTotal_ + (numDataTotal_ >> 2);
        hostAlloc((void**)&hData_, hDataCap_);
    }
    hipStream_t s = stream_;
    hipCheck(hipMemcpyAsync(hInfo_, dDenseInfo_, (size_t)nblocks_ * 20, hipMemcpyDeviceToHost, s), "result download");
    hipCheck(hipMemcpyAsync(hPassLen_, dDensePassLen_, numPassesTotal_ * 4, hipMemcpyDeviceToHost, s), "result download");
    hipCheck(hipMemcpyAsync(hPassDist_, dDensePassDist_, numPassesTotal_ * 4, hipMemcpyDeviceToHost, s), "result download");
    hipCheck(hipMemcpyAsync(hData_, dDenseData_, numDataTotal_, hipMemcpyDeviceToHost, s), "result download");
    if (anyRc_)
        hipCheck(hipMemcpyAsync(hRc_, dRc_, sizeof(RcFrame) * numFrames(), hipMemcpyDeviceToHost, s), "rate-control results");
    hipCheck(hipEventRecord(ev_[7], s), "event");
    hipCheck(hipStreamSynchronize(s), "result download");
    // frame-local offsets
    for (int f = 0; f < numFrames(); ++f) {
        uint32_t* inf = hInfo_ + (size_t)sd_[f].firstBlock * 5;
        const uint32_t d0 = inf[2], p0 = inf[4];
        frameDataBase_[f] = d0;
        framePassBase_[f] = p0;
        for (uint32_t b = 0; b < sd_[f].numBlocks; ++b) {
            inf[5 * b + 2] -= d0;
            inf[5 * b + 4] -= p0;
        }
    }
    {
        StageTimes t;
        auto ms = [&](int a, int b) {
            float v = 0;
            return hipEventElapsedTime(&v, ev_[a], ev_[b]) == hipSuccess ? v : 0.f;
        };
        if (!opt_.useGraph) {
            t.upload = ms(0, 1);
            t.color = ms(1, 2);
            t.dwt = ms(2, 3);
            t.dwt0 = ms(2, 8);
            t.t1cm = ms(3, 4);
            t.t1mq = ms(4, 5);
        }
        t.compact = opt_.useGraph ? 0.f : ms(5, 6);
        t.download = ms(6, 7);
        t.gpu = ms(0, 7);
        t.bytes = (double)numDataTotal_;
        t.symbols = (double)((const JobCounters*)hSmall_)->symbols;
        t.reruns = reruns;
        last_ = t;
        if (times) *times = t;
    }
}

void Job::encode(bool uploadInputs, StageTimes* times) {
    launch(uploadInputs);
    finish(times);
}

const RcFrame* Job::rateResult(int i) const {
    if (i < 0 || i >= numFrames() || !hRc_[i].max_total) return nullptr;
    return &hRc_[i];
}

Tier1View Job::result(int i) const {
    const SlotDev& d = sd_[i];
    const size_t next = i + 1 < numFrames() ? frameDataBase_[i + 1] : numDataTotal_;
    const size_t nextP = i + 1 < numFrames() ? framePassBase_[i + 1] : numPassesTotal_;
    Tier1View v;
    v.info5 = hInfo_ + (size_t)d.firstBlock * 5;
    v.passLen = hPassLen_ + framePassBase_[i];
    v.passDist = hPassDist_ + framePassBase_[i];
    v.data = hData_ + frameDataBase_[i];
    v.numBlocks = d.numBlocks;
    v.numPasses = nextP - framePassBase_[i];
    v.numData = next - frameDataBase_[i];
    return v;
}

void Job::debugCodes(int i, std::vector<uint16_t>& out) {
    out.resize(3 * sd_[i].npix);
    hipCheck(hipStreamSynchronize(stream_), "sync");
    hipCheck(hipMemcpy(out.data(), sd_[i].codes, out.size() * 2, hipMemcpyDeviceToHost), "codes download");
}

void Job::debugWavelet(int i, std::vector<float>& out) {
    // the padded planes, repacked into W x H Mallat planes
    const SlotDev& d = sd_[i];
    const size_t W = frames_[i].layout.width, H = frames_[i].layout.height, snx = (W + 1) / 2;
    std::vector<float> dev(3 * d.coefPlane);
    hipCheck(hipStreamSynchronize(stream_), "sync");
    hipCheck(hipMemcpy(dev.data(), d.coef, dev.size() * 4, hipMemcpyDeviceToHost), "wavelet download");
    out.resize(3 * d.npix);
    for (size_t c = 0; c < 3; ++c)
        for (size_t y = 0; y < H; ++y) {
            const float* src = dev.data() + c * d.coefPlane + y * d.coefStride;
            float* dst = out.data() + c * d.npix + y * W;
            std::memcpy(dst, src, snx * 4);
            std::memcpy(dst + snx, src + d.hiOff, (W - snx) * 4);
        }
}

void Job::debugCoeffs(int i, std::vector<uint32_t>& out) {
    const SlotDev& d = sd_[i];
    uint32_t* tmp = nullptr;
    hipCheck(hipMalloc(&tmp, (size_t)d.numBlocks * 4096), "debug buffer");
    hipCheck(launch_quant_debug(dBlocks_ + d.firstBlock, dBands_, dFrames_, d.numBlocks, tmp, stream_), "quantiser");
    out.resize((size_t)d.numBlocks * 1024);
    hipCheck(hipStreamSynchronize(stream_), "sync");
    const hipError_t e = hipMemcpy(out.data(), tmp, out.size() * 4, hipMemcpyDeviceToHost);
    hipFree(tmp);
    hipCheck(e, "coefficient download");
}

void Job::debugInfo(int i, std::vector<T1Info>& out) {
    const SlotDev& d = sd_[i];
    out.resize(d.numBlocks);
    hipCheck(hipStreamSynchronize(stream_), "sync");
    hipCheck(hipMemcpy(out.data(), dInfo_ + d.firstBlock, d.numBlocks * sizeof(T1Info), hipMemcpyDeviceToHost), "info");
}

}  // namespace gail_amd
