// Launch descriptors and launchers of the HIP decoder: Tier-1 block decoding, the inverse
// 9/7 DWT with dequantisation on load, and the inverse ICT with packing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gail_amd {

// One lane's code-block. Lanes left empty have dst == kDecEmpty.
struct DecBlock {
    uint32_t off, len;  // the codeword within the codestream buffer
    uint32_t geom;      // (w - 1) | (h - 1) << 5 | orient << 10 | nbps << 12 | passes << 18
    uint32_t dst;       // word index of the block's top-left sample in the Mallat planes
};
constexpr uint32_t kDecEmpty = 0xFFFFFFFFu;
// Only the first kDecWaveLanes lanes of a wave decode blocks (waves are sparse anyway: a wave
// runs as long as its longest lane, so long blocks go nearly alone), which lets a workgroup
// of kDecWaves waves keep every decoding lane's state in the LDS a full wave would take.
constexpr int kDecWaveLanes = 16;
constexpr int kDecWaves = 4;
constexpr int kDecSlots = kDecWaveLanes * kDecWaves;

struct T1DecodeParams {
    const uint8_t* bytes;  // codestream, readable 64 bytes past its end
    const DecBlock* blocks;  // nlong long blocks (one wave each), then kDecWaveLanes per wave
    uint32_t nblocks;      // entries, covering whole workgroups of kDecWaves waves
    uint32_t nlong;
    uint32_t* words;       // 3 Mallat planes of reconstruction words, zero beforehand
    uint32_t nwords;
    uint32_t stride;       // words per plane row (the frame width)
};

// One inverse DWT level: the region [0, w) x [0, h) of each component, its four quadrants
// (LL, HL, LH, HH) in Mallat order, back to interleaved samples.
struct IdwtLevel {
    const uint32_t* words;  // reconstruction words of the three components
    const float* ll;        // this level's LL quadrant from the level above; null: from words
    float* dst;             // output region (the next level's LL)
    uint32_t stride, plane; // elements per row / per component, same for all three buffers
    int32_t w, h;
    int32_t tiles_x, tiles;
    float half[3][4];       // half the quantisation step per component and orientation
};

// Dequantisation of a whole frame's words (the planes before the inverse DWT).
struct DequantFrame {
    const uint32_t* words;
    float* dst;
    uint32_t width, height, levels;
    float half[3][9][4];    // [component][level 1..8][orientation]
};

struct IctPack {
    const float* planes;    // three float planes, stride elements per row
    uint32_t stride, plane;
    int32_t w, h;           // picture size
    uint8_t* out;           // rows pitch bytes apart, three 16-bit words per pixel
    int64_t pitch;
    int32_t shift;          // 4: 12-bit codes in the high bits of each word, 0: in the low
};

hipError_t dec_tables_init();
// Waves (and so workgroups) the table of p covers.
inline uint32_t t1_decode_waves(uint32_t nblocks, uint32_t nlong) {
    return nlong + (nblocks - nlong) / (uint32_t)kDecWaveLanes;
}
hipError_t launch_t1_decode(const T1DecodeParams& p, hipStream_t s);
void idwt_tile_size(int& tx, int& ty);
hipError_t launch_idwt_level(const IdwtLevel& p, hipStream_t s);
hipError_t launch_dequant_frame(const DequantFrame& p, hipStream_t s);
hipError_t launch_ict_pack(const IctPack& p, hipStream_t s);
// Zeroes n 32-bit words at p (16-byte aligned) with a kernel rather than a memset node.
hipError_t launch_zero_words(uint32_t* p, size_t n, hipStream_t s);

}  // namespace gail_amd
