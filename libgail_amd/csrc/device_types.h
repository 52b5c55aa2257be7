// Plain structures shared by the host engine and the HIP kernels. Every table lives in
// HBM for the lifetime of a job; sizes are in the comments.
#pragma once
#include <cstdint>

namespace gail_amd {

constexpr int kMaxPasses = 64;       // pass table capacity per code-block
constexpr int kMaxPlanes = 24;       // magnitude bit-planes a code-block can carry
constexpr int kPlaneSymCap = 2560;   // symbols one bit-plane of a 32x32 block can produce
constexpr uint32_t kLog2Max = 2560;  // counts the entropy-estimate table covers

// One code-block of a job (16 bytes). coef_off is the float index of the block's
// top-left sample inside its frame's coefficient planes (comp * W * H + y * W + x).
struct T1Block {
    uint32_t coef_off;
    uint16_t frame;   // index into the job's frame table
    uint16_t band;    // index into the job's band table
    uint8_t w, h, orient, pad;
    uint32_t frame_block;  // index of the block within its frame
};

// One subband of one frame of the job (16 bytes).
struct T1Band {
    float inv_delta;    // 1 / delta, as the quantiser multiplies
    uint32_t max_mag;   // ((1 << Mb) << 6) - 1
    float dist_scale;   // delta^2 * weight: distortion units of the rate control
    uint32_t comp_mb;   // component | Mb << 8
};

// Per-frame device pointers of a job.
struct T1FrameRef {
    const float* coef;   // 3 Mallat planes, W x H each
    uint32_t stride;     // W
    float slope_thresh;  // early termination threshold of the frame (0: off)
};

// Per-block Tier-1 bookkeeping written by the context-modelling kernel.
struct T1Info {
    uint32_t nbps;     // significant magnitude bit-planes (0: nothing coded)
    uint32_t passes;   // coding passes produced
    uint32_t planes;   // bit-planes whose symbols were written
    uint32_t symbols;  // total symbols
};

// MQ coder state at the end of a pass (16 bytes).
struct MqSnapshot {
    int32_t bp;       // index of the pending byte (-1 before the first byte)
    uint32_t b_ct;    // pending byte | CT << 16
    uint32_t c, a;
};

// Counters of a job, cleared by a kernel before each run (never by a memset node).
struct JobCounters {
    unsigned long long sym_next;  // bump allocator of the symbol pool
    uint32_t flags;               // 1: symbol pool full, 2: codeword buffer full
    uint32_t pad;
    unsigned long long symbols;   // Tier-1 symbols of the job
    uint32_t pad2[10];
};

// ---------------------------------------------------------------- rate control on the device
constexpr int kRcMaxBins = 65536;  // slope histogram bins per component (64 octaves)

// One packet of a frame's codestream (8 bytes): precinct-bands [first, first + nband) and
// the tile-part it belongs to.
struct RcPacket {
    uint32_t first;
    uint16_t nband, tp;
};

// One precinct-band (12 bytes): the code-blocks block0 + j * stride + i (frame-local) for
// i < ncbx, j < ncby, of a subband with Mb magnitude bit-planes.
struct RcPBand {
    uint32_t block0;
    uint16_t stride;
    uint8_t ncbx, ncby, mb, pad[3];
};

// A point of a code-block's rate-distortion hull (24 bytes).
struct RcPoint {
    double slope, gain;
    uint32_t rate, passes;
};

// Rate control of one frame of a job: its parameters and, after the run, its result.
struct RcFrame {
    const RcPacket* packets;
    const RcPBand* pbands;
    uint32_t npackets, ntp;
    uint32_t first_block, nblocks;       // the frame's blocks in the job
    unsigned long long fixed_bytes;      // main header, COM reserve and EOC
    unsigned long long max_total, max_comp;  // budgets; max_total 0: no rate control
    int32_t tp_comp[6];                  // component of each tile-part
    // slope histogram bins [k_lo, k_lo + nb) of the run
    int64_t k_lo;
    int32_t nb, pad;
    // result
    double lambda[3];
    double min_slope, max_slope;
    unsigned long long total, comp[3], tile_part[6];
    double distortion;
    uint32_t iterations, status;         // status 0: done
};

}  // namespace gail_amd
