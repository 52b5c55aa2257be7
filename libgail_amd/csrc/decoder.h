// Decoder slot of the HIP backend: one frame's located code-blocks (what a Tier-2 reader
// hands over) to X'Y'Z' pixels on one device, the counterpart of the reference's CUDA
// decode slot (cuda_decode.cu:131-256). Header and packet parsing stay with the caller.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "kernels_dec.h"
#include "layout.h"

namespace gail_amd {

// gail::CblkCode (t2_parse.h:17-22): a code-block's bytes within the codestream, its coding
// passes (0: nothing coded) and its magnitude bit-planes.
struct DecCode {
    uint32_t offset, len;
    uint8_t passes, nbps, pad[2];
};

struct DecodeTimes {
    double t1 = 0, dwt = 0, color = 0, total = 0;  // ms; t1 includes the codestream upload
};

constexpr int kMaxDecLevels = 8;

// The Tier-1 decode table of one frame: the blocks to decode, longest first; the longest get
// a wave each, the rest are packed into waves of up to kDecWaveLanes blocks by a byte budget
// (a wave runs as long as its longest lane).
struct DescPlan {
    std::vector<uint32_t> order, count;
    uint32_t nlong = 0;
    uint64_t budget = 0;
    // wanted: the layout's blocks of the resolutions kept; coded[k]: wanted[k] has data.
    void plan(const Layout& L, const std::vector<int>& wanted, const uint8_t* coded, const DecCode* blocks, int cus);
    // Writes the table to out (nullptr: only counts); returns its entries (whole workgroups).
    uint32_t pack(const Layout& L, const DecCode* blocks, DecBlock* out) const;
};

// Device buffers of one frame's decode.
struct DecodeBuffers {
    const uint8_t* bytes = nullptr;   // codestream, readable 64 bytes past its end
    const DecBlock* descs = nullptr;  // nd table entries, the first nlong a wave each
    uint32_t nd = 0, nlong = 0;
    uint32_t* words = nullptr;        // 3 W x H reconstruction words, zero beforehand
    float* plane[2] = {nullptr, nullptr};  // 3 W x H floats each, the inverse DWT's ping-pong
    float* wave = nullptr;            // dequantised planes wanted (3 W x H floats), or null
    uint8_t* out = nullptr;           // the picture, rows pitch bytes apart
    int64_t pitch = 0;
};

// Queues the decode of one frame on s: Tier-1, dequantisation, the inverse DWT down to
// `reduce` levels left out, the inverse ICT into b.out. L: the geometry, frame: the frame's own
// step sizes. ev (may be null): two events, recorded after Tier-1 and after the inverse DWT.
void queueDecode(const Layout& L, const Layout& frame, int reduce, bool highBits, const DecodeBuffers& b, hipStream_t s,
                 hipEvent_t* ev);

// Decodes one frame with no device memory of its own: the intermediate buffers are a region of
// a fixed array in the plugin's device code (kNoAllocBytes per GPU, held until the process
// ends), for frames up to kNoAllocMaxWidth x kNoAllocMaxHeight with at most kNoAllocMaxBytes
// of codestream. data may be host or device memory; dst must be device memory. Synchronous:
// returns when the picture is in dst. Calls on one GPU share the array; the library takes
// them one at a time.
constexpr int kNoAllocMaxWidth = 2048, kNoAllocMaxHeight = 1080;
constexpr size_t kNoAllocMaxBytes = 1536 * 1024;
void decodeNoAlloc(int device, const Layout& frame, int reduce, bool highBits, const uint8_t* data, size_t size,
                   const DecCode* blocks, size_t numBlocks, void* dst, ptrdiff_t strideBytes, hipStream_t stream);

class Decoder {
public:
    // reduce: resolution levels left out (0..levels); highBits: codes << 4 in each word.
    Decoder(int device, const Layout& L, int reduce, bool highBits);
    ~Decoder();
    Decoder(const Decoder&) = delete;
    Decoder& operator=(const Decoder&) = delete;

    // Device memory a decoder for L takes (gpuDecoderSlotBytes, decode_backend.h:58-63,
    // with this decoder's buffers).
    static uint64_t bytesFor(const Layout& L, int reduce);

    // frame: the layout of this frame (the geometry the decoder was made for, with the step
    // sizes of the frame's own headers). data: the codestream (host or device memory) the
    // blocks point into; blocks: one per code-block of the layout. dst: outWidth() x
    // outHeight() pixels of three 16-bit words, rows strideBytes apart, in host or device
    // memory. wantWavelet keeps the dequantised planes for wavelet().
    void decode(const Layout& frame, const uint8_t* data, size_t size, const DecCode* blocks, size_t n, void* dst,
                ptrdiff_t strideBytes, bool wantWavelet, DecodeTimes* tm);

    int outWidth() const { return outW_; }
    int outHeight() const { return outH_; }
    // Plane c of the dequantised coefficients of the last decode(wantWavelet) (W x H floats).
    const float* wavelet(int c) const { return hWave_.data() + (size_t)c * L_.width * L_.height; }
    size_t waveletSamples() const { return (size_t)L_.width * L_.height; }

private:
    void release();
    void ensureBytes(size_t n);
    void ensureDescs(size_t n);
    // Orders the blocks marked in coded[] (one flag per wanted block) longest first, gives the
    // longest a wave each and packs the rest into waves; returns the table's entries.
    uint32_t buildDescs(const uint8_t* coded, const DecCode* blocks);

    int dev_;
    Layout L_;
    int reduce_;
    bool highBits_;
    int outW_ = 0, outH_ = 0;
    std::vector<int> ws_, hs_, wanted_;
    DescPlan plan_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[4] = {};
    uint32_t* dWords_ = nullptr;
    float* dPlane_[2] = {};
    float* dWave_ = nullptr;
    uint16_t* dOut_ = nullptr;
    uint16_t* hOut_ = nullptr;
    uint8_t* dBytes_ = nullptr;
    uint8_t* hBytes_ = nullptr;
    size_t capBytes_ = 0;
    DecBlock* dDescs_ = nullptr;
    DecBlock* hDescs_ = nullptr;
    size_t capDescs_ = 0;
    uint32_t nlong_ = 0;  // leading table entries decoded a whole wave each
    int cus_ = 256;
    std::vector<float> hWave_;
};

}  // namespace gail_amd
