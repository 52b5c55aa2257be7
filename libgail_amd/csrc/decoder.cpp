// Decoder slot of the HIP backend (decoder.h). Host code compiled by hipcc.
#include "decoder.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.h"
#include "errors.h"

namespace gail_amd {

namespace {

// Bytes a lane may read past a codeword's end (the 16-byte window, aligned down to 8).
constexpr size_t kBytePad = 64;
// Length classes of the block ordering (bytes / 4, longer blocks share the top class).
constexpr uint32_t kLenKeys = 4096;

enum class Mem { Pageable, Pinned, Device };

// Where a buffer lives: device memory is read or written in place, pinned host memory by DMA,
// anything else through the decoder's own pinned staging buffers.
Mem memKind(const void* p) {
    hipPointerAttribute_t a{};
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return Mem::Pageable;
    }
    if (a.type == hipMemoryTypeDevice) return Mem::Device;
    if (a.type == hipMemoryTypeHost) return Mem::Pinned;
    return Mem::Pageable;
}

double msSince(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

uint64_t Decoder::bytesFor(const Layout& L, int reduce) {
    int w = L.width, h = L.height;
    for (int i = 0; i < reduce; ++i) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    const uint64_t npix = (uint64_t)L.width * L.height;
    // words + two float plane sets + the packed picture + codestream and block table room
    return 3 * npix * 4 * 3 + (uint64_t)w * h * 6 + (4u << 20);
}

Decoder::Decoder(int device, const Layout& L, int reduce, bool highBits)
    : dev_(device), L_(L), reduce_(reduce), highBits_(highBits) {
    if (reduce < 0 || reduce > L.levels) throw std::runtime_error("more resolution levels left out than the stream has");
    try {
        hipCheck(hipSetDevice(dev_), "hipSetDevice");
        cus_ = DeviceContext::get(dev_).computeUnits();
        ws_.resize(L_.levels + 1);
        hs_.resize(L_.levels + 1);
        ws_[0] = L_.width;
        hs_[0] = L_.height;
        for (int l = 1; l <= L_.levels; ++l) {
            ws_[l] = (ws_[l - 1] + 1) / 2;
            hs_[l] = (hs_[l - 1] + 1) / 2;
        }
        outW_ = ws_[reduce];
        outH_ = hs_[reduce];
        for (int i = 0; i < L_.numCblks(); ++i)
            if (L_.bands[L_.cblks[i].band].res <= L_.levels - reduce) wanted_.push_back(i);
        hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
        for (auto& e : ev_) hipCheck(hipEventCreate(&e), "event");
        const size_t npix = (size_t)L_.width * L_.height;
        hipCheck(hipMalloc(&dWords_, 3 * npix * sizeof(uint32_t)), "decoder words");
        hipCheck(hipMemset(dWords_, 0, 3 * npix * sizeof(uint32_t)), "decoder words");
        for (auto& p : dPlane_) hipCheck(hipMalloc(&p, 3 * npix * sizeof(float)), "decoder planes");
        const size_t outBytes = (size_t)outW_ * outH_ * 6;
        hipCheck(hipMalloc(&dOut_, outBytes), "decoder picture");
        hipCheck(hipHostMalloc((void**)&hOut_, outBytes, hipHostMallocDefault), "decoder picture (pinned)");
        ensureBytes(1500000);  // a DCI frame is at most 250 Mbit/s over 24 frames
        ensureDescs(((size_t)wanted_.size() / kDecSlots + 1) * kDecSlots);
    } catch (...) {
        release();
        throw;
    }
}

Decoder::~Decoder() { release(); }

void Decoder::release() {
    (void)hipSetDevice(dev_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : {(void*)dWords_, (void*)dPlane_[0], (void*)dPlane_[1], (void*)dWave_, (void*)dOut_, (void*)dBytes_,
                    (void*)dDescs_})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)hOut_, (void*)hBytes_, (void*)hDescs_})
        if (p) (void)hipHostFree(p);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
    dWords_ = nullptr;
    dPlane_[0] = dPlane_[1] = nullptr;
    dWave_ = nullptr;
    dOut_ = hOut_ = nullptr;
    dBytes_ = hBytes_ = nullptr;
    dDescs_ = hDescs_ = nullptr;
    stream_ = nullptr;
    (void)hipGetLastError();
}

void Decoder::ensureBytes(size_t n) {
    if (n <= capBytes_) return;
    if (dBytes_) (void)hipFree(dBytes_);
    if (hBytes_) (void)hipHostFree(hBytes_);
    dBytes_ = hBytes_ = nullptr;
    capBytes_ = 0;
    const size_t cap = n + n / 2;
    hipCheck(hipMalloc(&dBytes_, cap + kBytePad), "codestream buffer");
    hipCheck(hipMemset(dBytes_, 0, cap + kBytePad), "codestream buffer");
    hipCheck(hipHostMalloc((void**)&hBytes_, cap, hipHostMallocDefault), "codestream buffer (pinned)");
    capBytes_ = cap;
}

void Decoder::ensureDescs(size_t n) {
    if (n <= capDescs_) return;
    if (dDescs_) (void)hipFree(dDescs_);
    if (hDescs_) (void)hipHostFree(hDescs_);
    dDescs_ = hDescs_ = nullptr;
    capDescs_ = 0;
    const size_t cap = n + n / 2;
    hipCheck(hipMalloc(&dDescs_, cap * sizeof(DecBlock)), "block table");
    hipCheck(hipHostMalloc((void**)&hDescs_, cap * sizeof(DecBlock), hipHostMallocDefault), "block table (pinned)");
    capDescs_ = cap;
}

void DescPlan::plan(const Layout& L, const std::vector<int>& wanted, const uint8_t* coded, const DecCode* blocks,
                    int cus) {
    // Coded blocks longest first (a counting sort over length classes).
    count.assign(kLenKeys + 1, 0u);
    uint64_t total = 0;
    size_t ncoded = 0;
    auto key = [](uint32_t len) { return kLenKeys - 1u - std::min(len >> 2, kLenKeys - 1u); };
    for (size_t k = 0; k < wanted.size(); ++k)
        if (coded[k]) {
            ++count[key(blocks[wanted[k]].len) + 1];
            total += blocks[wanted[k]].len;
            ++ncoded;
        }
    for (uint32_t k = 0; k < kLenKeys; ++k) count[k + 1] += count[k];
    order.resize(ncoded);
    for (size_t k = 0; k < wanted.size(); ++k)
        if (coded[k]) order[count[key(blocks[wanted[k]].len)]++] = (uint32_t)wanted[k];
    // Waves: a wave runs as long as its longest lane, so a long block takes a wave of its own
    // and short ones share, up to kDecWaveLanes blocks and a byte budget that spreads the frame
    // over about eight waves per compute unit (measured best on 4K frames at 250 Mbit/s).
    budget = std::max<uint64_t>(384, total / (uint64_t)(8 * cus));
    // blocks over the budget would sit alone in a wave: they get a whole wave each and are
    // decoded with the wave's scalar unit
    nlong = 0;
    while (nlong < order.size() && blocks[order[nlong]].len > budget) ++nlong;
    (void)L;
}

uint32_t DescPlan::pack(const Layout& L, const DecCode* blocks, DecBlock* out) const {
    const uint32_t lanes = (uint32_t)kDecWaveLanes;
    auto entry = [&](uint32_t i) {
        const DecCode& c = blocks[i];
        const CodeBlock& cb = L.cblks[i];
        const Subband& b = L.bands[cb.band];
        DecBlock d;
        d.off = c.offset;
        d.len = c.len;
        d.geom = (uint32_t)(cb.w - 1) | ((uint32_t)(cb.h - 1) << 5) | ((uint32_t)b.orient << 10) |
                 ((uint32_t)c.nbps << 12) | ((uint32_t)c.passes << 18);
        d.dst = (uint32_t)b.comp * (uint32_t)L.width * (uint32_t)L.height +
                (uint32_t)(b.py0 + cb.y0) * (uint32_t)L.width + (uint32_t)(b.px0 + cb.x0);
        return d;
    };
    uint32_t n = 0, inWave = 0;
    uint64_t waveBytes = 0;
    for (uint32_t k = 0; k < nlong; ++k, ++n)
        if (out) out[n] = entry(order[k]);
    for (uint32_t k = nlong; k < order.size(); ++k) {
        const uint32_t i = order[k];
        const DecCode& c = blocks[i];
        if (inWave && (inWave == lanes || waveBytes + c.len > budget)) {
            if (out)
                for (; inWave < lanes; ++inWave) out[n++] = DecBlock{0, 0, 0, kDecEmpty};
            else
                n += lanes - inWave;
            inWave = 0;
            waveBytes = 0;
        }
        if (out) out[n] = entry(i);
        ++n;
        ++inWave;
        waveBytes += c.len;
    }
    if (inWave) {
        if (out)
            for (; inWave < lanes; ++inWave) out[n++] = DecBlock{0, 0, 0, kDecEmpty};
        else
            n += lanes - inWave;
    }
    // whole workgroups
    while (t1_decode_waves(n, nlong) % (uint32_t)kDecWaves) {
        for (uint32_t j = 0; j < lanes; ++j, ++n)
            if (out) out[n] = DecBlock{0, 0, 0, kDecEmpty};
    }
    return n;
}

uint32_t Decoder::buildDescs(const uint8_t* coded, const DecCode* blocks) {
    plan_.plan(L_, wanted_, coded, blocks, cus_);
    const uint32_t n = plan_.pack(L_, blocks, nullptr);
    ensureDescs(n);
    plan_.pack(L_, blocks, hDescs_);
    nlong_ = plan_.nlong;
    return n;
}

void queueDecode(const Layout& L, const Layout& frame, int reduce, bool highBits, const DecodeBuffers& b, hipStream_t s,
                 hipEvent_t* ev) {
    const uint32_t W = (uint32_t)L.width, H = (uint32_t)L.height;
    const size_t npix = (size_t)W * H;
    T1DecodeParams tp;
    tp.bytes = b.bytes;
    tp.blocks = b.descs;
    tp.nblocks = b.nd;
    tp.nlong = b.nlong;
    tp.words = b.words;
    tp.nwords = (uint32_t)(3 * npix);
    tp.stride = W;
    hipCheck(launch_t1_decode(tp, s), "Tier-1 decode");

    // half the step of every band, from this frame's own step sizes
    auto half = [&](int comp, int level, int orient) {
        const int res = orient == 0 ? 0 : L.levels - level + 1;
        return 0.5f * frame.bands[L.bandIndex(comp, res, orient)].delta;
    };
    if (b.wave || reduce == L.levels) {
        DequantFrame dq{};
        dq.words = b.words;
        dq.dst = b.wave ? b.wave : b.plane[0];
        dq.width = W;
        dq.height = H;
        dq.levels = (uint32_t)L.levels;
        for (int c = 0; c < 3; ++c)
            for (int l = 1; l <= L.levels; ++l)
                for (int o = 0; o < 4; ++o) dq.half[c][l][o] = (o == 0 && l != L.levels) ? 0.f : half(c, l, o);
        hipCheck(launch_dequant_frame(dq, s), "dequantisation");
        if (reduce == L.levels && b.wave) {
            dq.dst = b.plane[0];
            hipCheck(launch_dequant_frame(dq, s), "dequantisation");
        }
    }
    if (ev) hipCheck(hipEventRecord(ev[0], s), "event");

    // inverse DWT, the top level first; each level's output is the next one's LL
    int ws[kMaxDecLevels + 1], hs[kMaxDecLevels + 1];
    ws[0] = L.width;
    hs[0] = L.height;
    for (int l = 1; l <= L.levels; ++l) {
        ws[l] = (ws[l - 1] + 1) / 2;
        hs[l] = (hs[l - 1] + 1) / 2;
    }
    const float* ll = reduce == L.levels ? b.plane[0] : nullptr;
    int cur = 0, tx, ty;
    idwt_tile_size(tx, ty);
    for (int l = L.levels; l > reduce; --l) {
        IdwtLevel p{};
        p.words = b.words;
        p.ll = ll;
        p.dst = b.plane[cur];
        p.stride = W;
        p.plane = (uint32_t)npix;
        p.w = ws[l - 1];
        p.h = hs[l - 1];
        p.tiles_x = (p.w + tx - 1) / tx;
        p.tiles = p.tiles_x * ((p.h + ty - 1) / ty);
        for (int c = 0; c < 3; ++c)
            for (int o = 0; o < 4; ++o) p.half[c][o] = (o == 0 && l != L.levels) ? 0.f : half(c, l, o);
        hipCheck(launch_idwt_level(p, s), "inverse DWT");
        ll = b.plane[cur];
        cur ^= 1;
    }
    if (ev) hipCheck(hipEventRecord(ev[1], s), "event");

    IctPack ip;
    ip.planes = ll;
    ip.stride = W;
    ip.plane = (uint32_t)npix;
    ip.w = ws[reduce];
    ip.h = hs[reduce];
    ip.out = b.out;
    ip.pitch = b.pitch;
    ip.shift = highBits ? 4 : 0;
    hipCheck(launch_ict_pack(ip, s), "inverse ICT");
}

void Decoder::decode(const Layout& frame, const uint8_t* data, size_t size, const DecCode* blocks, size_t n, void* dst,
                     ptrdiff_t strideBytes, bool wantWavelet, DecodeTimes* tm) {
    if (frame.width != L_.width || frame.height != L_.height || frame.levels != L_.levels ||
        frame.bands.size() != L_.bands.size() || frame.numCblks() != L_.numCblks() || n != (size_t)L_.numCblks())
        throw std::logic_error("frame does not match the layout the decoder was made for");
    if (!dst || (size && !data) || (n && !blocks)) throw std::logic_error("null codestream, block table or picture");
    if (size > 0xFFFFFFFFull - kBytePad) throw std::runtime_error("codestream too large");
    const ptrdiff_t rowBytes = (ptrdiff_t)outW_ * 6;
    if (strideBytes < rowBytes && strideBytes > -rowBytes) throw std::logic_error("row stride shorter than a picture row");
    const auto t0 = std::chrono::steady_clock::now();
    hipCheck(hipSetDevice(dev_), "hipSetDevice");
    const hipStream_t s = stream_;
    const uint32_t W = (uint32_t)L_.width, H = (uint32_t)L_.height;
    const size_t npix = (size_t)W * H;

    // Blocks decoded: of the resolutions kept, with passes, bit-planes the block coder carries
    // and bytes inside the codestream; the rest stay zero.
    std::vector<uint8_t> coded(wanted_.size());
    for (size_t k = 0; k < wanted_.size(); ++k) {
        const DecCode& c = blocks[wanted_[k]];
        coded[k] = c.passes && c.nbps >= 1 && c.nbps <= 30 && (size_t)c.offset + c.len <= size;
    }
    const uint32_t nd = buildDescs(coded.data(), blocks);
    const Mem in = memKind(data), outKind = memKind(dst);
    const bool devOut = outKind == Mem::Device;
    // pinned pictures are written by DMA when their rows go forwards
    const bool dmaOut = outKind == Mem::Pinned && strideBytes > 0;
    ensureBytes(std::max<size_t>(size, 1));
    hipCheck(hipEventRecord(ev_[0], s), "event");
    if (size) {
        if (in == Mem::Device) {
            hipCheck(hipMemcpyAsync(dBytes_, data, size, hipMemcpyDeviceToDevice, s), "codestream copy");
        } else if (in == Mem::Pinned) {
            hipCheck(hipMemcpyAsync(dBytes_, data, size, hipMemcpyHostToDevice, s), "codestream upload");
        } else {
            std::memcpy(hBytes_, data, size);
            hipCheck(hipMemcpyAsync(dBytes_, hBytes_, size, hipMemcpyHostToDevice, s), "codestream upload");
        }
    }
    if (nd) hipCheck(hipMemcpyAsync(dDescs_, hDescs_, (size_t)nd * sizeof(DecBlock), hipMemcpyHostToDevice, s), "block table upload");
    hipCheck(hipMemsetAsync(dWords_, 0, 3 * npix * sizeof(uint32_t), s), "clear words");
    if (wantWavelet && !dWave_) hipCheck(hipMalloc(&dWave_, 3 * npix * sizeof(float)), "wavelet planes");
    DecodeBuffers buf;
    buf.bytes = dBytes_;
    buf.descs = dDescs_;
    buf.nd = nd;
    buf.nlong = nlong_;
    buf.words = dWords_;
    buf.plane[0] = dPlane_[0];
    buf.plane[1] = dPlane_[1];
    buf.wave = wantWavelet ? dWave_ : nullptr;
    buf.out = devOut ? (uint8_t*)dst : (uint8_t*)dOut_;
    buf.pitch = devOut ? (int64_t)strideBytes : (int64_t)rowBytes;
    queueDecode(L_, frame, reduce_, highBits_, buf, s, &ev_[1]);
    const size_t outBytes = (size_t)outW_ * outH_ * 6;
    if (dmaOut) {
        hipCheck(hipMemcpy2DAsync(dst, (size_t)strideBytes, dOut_, (size_t)rowBytes, (size_t)rowBytes, (size_t)outH_,
                                  hipMemcpyDeviceToHost, s),
                 "picture download");
    } else if (!devOut) {
        hipCheck(hipMemcpyAsync(hOut_, dOut_, outBytes, hipMemcpyDeviceToHost, s), "picture download");
    }
    hipCheck(hipEventRecord(ev_[3], s), "event");
    if (wantWavelet) {
        hWave_.resize(3 * npix);
        hipCheck(hipMemcpyAsync(hWave_.data(), dWave_, 3 * npix * sizeof(float), hipMemcpyDeviceToHost, s), "wavelet download");
    }
    hipCheck(hipStreamSynchronize(s), "decode");
    hipCheck(hipGetLastError(), "decode");
    if (!devOut && !dmaOut) {
        if (strideBytes == rowBytes) {
            std::memcpy(dst, hOut_, outBytes);
        } else {
            for (int y = 0; y < outH_; ++y)
                std::memcpy((uint8_t*)dst + (ptrdiff_t)y * strideBytes, (const uint8_t*)hOut_ + (size_t)y * rowBytes, rowBytes);
        }
    }
    if (tm) {
        float a = 0, b = 0, c = 0;
        (void)hipEventElapsedTime(&a, ev_[0], ev_[1]);
        (void)hipEventElapsedTime(&b, ev_[1], ev_[2]);
        (void)hipEventElapsedTime(&c, ev_[2], ev_[3]);
        tm->t1 = a;
        tm->dwt = b;
        tm->color = c;
        tm->total = msSince(t0);
    }
}

// ---------------------------------------------------------------- decoding without allocation
namespace {

// Intermediate buffers of one frame of up to kNoAllocMaxWidth x kNoAllocMaxHeight: the
// reconstruction words and the two float plane sets (3 x 2048 x 1080 x 4 bytes each), the
// codestream with its read-ahead pad, and the Tier-1 table (up to 16 entries per block when
// every wave holds one block, for the 6,708 blocks of a 2K frame).
constexpr size_t kNoAllocPlane = (size_t)3 * kNoAllocMaxWidth * kNoAllocMaxHeight * 4;
constexpr size_t kNoAllocCodeBytes = (kNoAllocMaxBytes + kBytePad + 255) / 256 * 256;
constexpr size_t kNoAllocTableBytes = 2u << 20;
constexpr size_t kNoAllocBytes = 3 * kNoAllocPlane + kNoAllocCodeBytes + kNoAllocTableBytes;

// The array every GPU holds (in this file's code object) until the process ends. Host code
// uses its per-device address from hipGetSymbolAddress, never the name.
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;

struct NoAllocDevice {
    std::once_flag once;
    hipError_t setup = hipSuccess;
    uint8_t* base = nullptr;  // scratch's address on this device
    int cus = 256;
    std::mutex m;             // one call at a time on the shared array
};
NoAllocDevice g_noAlloc[kNoAllocDevices];

// With the device current: the array's address, the decoder's constant tables and the
// compute-unit count, once per device (no allocation: the tables go to constant memory).
void noAllocSetup(NoAllocDevice& d, int device) {
    std::call_once(d.once, [&] {
        void* p = nullptr;
        d.setup = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d.setup == hipSuccess) d.base = static_cast<uint8_t*>(p);
        if (d.setup == hipSuccess) d.setup = dec_tables_init();
        int cus = 0;
        if (d.setup == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
            cus > 0)
            d.cus = cus;
    });
    hipCheck(d.setup, "no-allocation decoder set-up");
}

}  // namespace

void decodeNoAlloc(int device, const Layout& frame, int reduce, bool highBits, const uint8_t* data, size_t size,
                   const DecCode* blocks, size_t numBlocks, void* dst, ptrdiff_t strideBytes, hipStream_t stream) {
    if (device < 0 || device >= kNoAllocDevices) throw std::logic_error("no such device");
    if (frame.width > kNoAllocMaxWidth || frame.height > kNoAllocMaxHeight)
        throw std::logic_error("a " + std::to_string(frame.width) + "x" + std::to_string(frame.height) +
                               " frame exceeds the no-allocation decoder's " + std::to_string(kNoAllocMaxWidth) + "x" +
                               std::to_string(kNoAllocMaxHeight) + " limit");
    if (size > kNoAllocMaxBytes)
        throw std::logic_error("a codestream of " + std::to_string(size) + " bytes exceeds the no-allocation decoder's " +
                               std::to_string(kNoAllocMaxBytes) + "-byte limit");
    if (frame.levels < 1 || frame.levels > kMaxDecLevels || reduce < 0 || reduce > frame.levels)
        throw std::logic_error("more resolution levels left out than the stream has");
    if (numBlocks != (size_t)frame.numCblks()) throw std::logic_error("one block record per code-block is needed");
    if (!dst || (size && !data) || (numBlocks && !blocks)) throw std::logic_error("null codestream, block table or picture");
    int outW = frame.width, outH = frame.height;
    for (int i = 0; i < reduce; ++i) {
        outW = (outW + 1) / 2;
        outH = (outH + 1) / 2;
    }
    const ptrdiff_t rowBytes = (ptrdiff_t)outW * 6;
    if (strideBytes < rowBytes && strideBytes > -rowBytes) throw std::logic_error("row stride shorter than a picture row");
    hipCheck(hipSetDevice(device), "hipSetDevice");
    if (memKind(dst) != Mem::Device) throw std::logic_error("the no-allocation decoder writes device memory only");
    NoAllocDevice& d = g_noAlloc[device];
    noAllocSetup(d, device);
    std::lock_guard<std::mutex> lock(d.m);

    // the table on the host, in storage each thread keeps from call to call
    thread_local std::vector<int> wanted;
    thread_local std::vector<uint8_t> coded;
    thread_local DescPlan plan;
    thread_local std::vector<DecBlock> table;
    wanted.clear();
    coded.clear();
    for (int i = 0; i < frame.numCblks(); ++i)
        if (frame.bands[frame.cblks[i].band].res <= frame.levels - reduce) {
            const DecCode& c = blocks[i];
            wanted.push_back(i);
            coded.push_back(c.passes && c.nbps >= 1 && c.nbps <= 30 && (size_t)c.offset + c.len <= size);
        }
    plan.plan(frame, wanted, coded.data(), blocks, d.cus);
    const uint32_t nd = plan.pack(frame, blocks, nullptr);
    if ((size_t)nd * sizeof(DecBlock) > kNoAllocTableBytes) throw std::runtime_error("Tier-1 table larger than its room");
    table.resize(nd);
    plan.pack(frame, blocks, table.data());

    // the array's regions
    const size_t npix = (size_t)frame.width * frame.height;
    DecodeBuffers b;
    b.words = reinterpret_cast<uint32_t*>(d.base);
    b.plane[0] = reinterpret_cast<float*>(d.base + kNoAllocPlane);
    b.plane[1] = reinterpret_cast<float*>(d.base + 2 * kNoAllocPlane);
    uint8_t* bytes = d.base + 3 * kNoAllocPlane;
    DecBlock* descs = reinterpret_cast<DecBlock*>(bytes + kNoAllocCodeBytes);
    b.bytes = bytes;
    b.descs = descs;
    b.nd = nd;
    b.nlong = plan.nlong;
    b.out = static_cast<uint8_t*>(dst);
    b.pitch = (int64_t)strideBytes;

    const hipStream_t s = stream;
    if (size) hipCheck(hipMemcpyAsync(bytes, data, size, hipMemcpyDefault, s), "codestream copy");
    if (nd) hipCheck(hipMemcpyAsync(descs, table.data(), (size_t)nd * sizeof(DecBlock), hipMemcpyHostToDevice, s), "table upload");
    // the words start at zero: the last call left its own in the array
    hipCheck(launch_zero_words(b.words, 3 * npix, s), "clear words");
    queueDecode(frame, frame, reduce, highBits, b, s, nullptr);
    hipCheck(hipStreamSynchronize(s), "decode");
    hipCheck(hipGetLastError(), "decode");
}

}  // namespace gail_amd
