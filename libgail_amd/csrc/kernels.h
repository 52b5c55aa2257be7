// Kernel launch descriptors and launchers of the HIP encoder backend.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_types.h"

namespace gail_amd {

// One frame's part of a DWT level launch. Tiles of all frames are numbered
// consecutively; tile_base is this frame's first.
struct DwtLevelDesc {
    const void* src;        // level 0: 3 planes of u16 codes; else 3 float LL planes
    uint32_t src_stride;    // elements per row
    uint32_t src_plane;     // elements per plane
    float* coef;            // coefficient planes of the frame (Mallat order, rows padded)
    uint32_t coef_stride, coef_plane;
    uint32_t hi_off;        // column of this level's high band in coef (level 0: padded)
    float* ll_dst;          // where this level's LL quarter goes
    uint32_t ll_stride, ll_plane;
    int32_t w, h;           // LL region transformed by this level
    int32_t tiles_x;
    uint32_t tile_base;
};

// Colour conversion of one frame to planar 12-bit X'Y'Z' codes.
struct ColorDesc {
    const void* src;
    uint16_t* codes;          // 3 planes of width * height
    const float* lut;         // 65,536-entry EOTF table, cd/m^2
    const float* thresholds;  // 4,096 code thresholds in cd/m^2
    const float* lut3;        // app 3D LUT (lut3n^3 x 3 floats, red fastest) when lut3n > 1
    int32_t lut3n;
    float params[48];         // colour parameter block (RGB -> XYZ, tone map, gamut)
    int32_t width, height;
    int32_t kind;             // 0 interleaved 16-bit R'G'B', 1 planar
    int32_t aw, ah, ox, oy, sx, sy, bits, rgb, full_range, chroma;
    float crR, cbG, crG, cbB;
};

struct CmParams {
    const T1Block* blocks;
    const T1Band* bands;
    const T1FrameRef* frames;
    uint32_t nblocks;
    uint32_t max_passes;
    T1Info* info;
    uint32_t* pass_end;   // [block][64] cumulative symbol counts
    int32_t* pass_dist;   // [block][64]
    uint2* chunks;        // [block][24]: (offset / 16, count) of each bit-plane's symbols
    uint8_t* pool;
    unsigned long long pool_cap;
    JobCounters* counters;
    uint32_t* work_next;  // persistent work counter (block pairs)
    const float* log2tab; // host log2f over the entropy estimate's domain
};

struct MqParams {
    const T1Info* info;
    const uint32_t* pass_end;
    const uint2* chunks;
    const uint8_t* pool;
    const uint32_t* perm;
    const unsigned long long* out_off;  // [block] codeword offset (exclusive scan of capacities)
    uint8_t* out;
    unsigned long long out_cap;
    MqSnapshot* snaps;                  // [block][64]
    uint32_t* pass_len;                 // [block][64]
    uint32_t* code_len;                 // [block] codeword length
    JobCounters* counters;
    uint32_t nblocks;
    uint32_t* work_next;
    uint32_t nlong;                     // leading blocks of perm coded one per wave (<= nblocks)
};

struct CompactParams {
    const T1Info* info;
    const uint32_t* code_len;
    const unsigned long long* out_off;
    const uint8_t* out;
    const uint32_t* pass_len;
    const int32_t* pass_dist;
    const unsigned long long* data_off;  // [block] exclusive scan of 4-byte-rounded lengths
    const uint32_t* pass_off;            // [block] exclusive scan of pass counts
    uint32_t* dense_info;                // [block][5]: nbps, passes, data offset, length, pass base
    uint32_t* dense_pass_len;
    int32_t* dense_pass_dist;
    uint8_t* dense_data;
    const JobCounters* counters;
    uint32_t nblocks;
};

// Device rate control (kernels_rc.hip), after MQ, before compaction: a hull launch over
// every block, then one workgroup per frame.
struct RcParams {
    RcFrame* frames;
    const T1Block* blocks;
    const T1Band* bands;
    const double* band_scale;   // [band] delta^2 * weight in double
    T1Info* info;               // passes overwritten with the chosen ones
    const uint32_t* pass_len;
    const int32_t* pass_dist;
    uint32_t* code_len;         // overwritten with the chosen passes' length
    // hull points, point-major: point h of block b at [h * stride + b]
    size_t stride;              // the job's block capacity
    double* h_slope;
    double* h_gain;
    uint32_t* h_sel;            // passes << 24 | truncation length
    int32_t* h_kbin;            // floor(1024 log2 slope)
    uint32_t* hull_n;           // [block] hull points | component << 8
    unsigned long long* part;   // [2 * rcPartEntries] slope range of each hull workgroup
    uint32_t* hist;             // [frame][6][kRcMaxBins + 1], zero between runs
    uint32_t* segsum;           // [rc_segsum_entries] scan segment totals
    const unsigned long long* mantissa;  // [1025] the octave's steps
    const uint16_t* first;               // [1024]
    uint32_t* sel;              // [block] chosen passes << 24 | their length
    const JobCounters* counters;
};
// Entries of RcParams::part for a job of `blocks` code-blocks in `frames` frames.
inline size_t rcPartEntries(size_t blocks, size_t frames) { return blocks / 256 + frames + 2; }
size_t rc_segsum_entries(size_t frames);
hipError_t launch_rc(const RcParams& p, int frames, uint32_t max_frame_blocks, hipStream_t s);

hipError_t launch_dwt_level(bool level0, int tile_shape, const DwtLevelDesc* descs, int ndesc, int tiles,
                            hipStream_t s);
// Tile shapes: 0 64x64, 1 64x32, 2 32x64, 3 64x16 (LDS tiles, 256 threads each), 4 120x64,
// 5 120x32, 6 120x128 (register-streaming strips, one wave each); 7, 8 level 0 as strips whose
// input rows are loaded ahead (7: 120x16, 10 rows ahead; 8: 120x8, every row up front) and the
// other levels as 64x16 LDS tiles. Levels whose LL region is narrower or lower than 2 samples
// take the LDS tiles.
constexpr int kDwtShapes = 11;
constexpr int kDwtDefaultShape = 10;
void dwt_tile_size(int tile_shape, int& tx, int& ty);
hipError_t launch_color(const ColorDesc& d, hipStream_t s);

// Uploads the constant context and probability tables; call once per device before any launch.
hipError_t t1_tables_init();
hipError_t launch_clear_counters(JobCounters* c, uint32_t* work, int nwork, hipStream_t s);
hipError_t launch_t1_cm(const CmParams& p, int waves_per_group, int groups, hipStream_t s);
hipError_t launch_sort_blocks(const T1Info* info, uint32_t n, uint32_t* hist, uint32_t* perm, JobCounters* counters,
                              hipStream_t s);
hipError_t launch_out_offsets(const T1Info* info, uint32_t n, unsigned long long* out_off, unsigned long long* partial,
                              hipStream_t s);
// group: lanes per workgroup (64, 128, 256); read_bytes: symbols read per lane at a time (16 or 64).
hipError_t launch_t1_mq(const MqParams& p, int group, int read_bytes, int groups, hipStream_t s);
hipError_t launch_compact_offsets(const T1Info* info, const uint32_t* code_len, uint32_t n,
                                  unsigned long long* data_off, uint32_t* pass_off, unsigned long long* partial,
                                  hipStream_t s);
hipError_t launch_compact(const CompactParams& p, hipStream_t s);
hipError_t launch_quant_debug(const T1Block* blocks, const T1Band* bands, const T1FrameRef* frames, uint32_t n,
                              uint32_t* out, hipStream_t s);

}  // namespace gail_amd
