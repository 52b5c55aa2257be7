// Device engine of the HIP backend: plans a job (one frame, or many frames of mixed
// geometry), owns its HBM buffers and stream, and runs colour -> DWT -> fused quantisation
// + context modelling -> ordering -> MQ -> compaction on it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "color_tables.h"
#include "device_types.h"
#include "kernels.h"
#include "layout.h"

namespace gail_amd {

// How frames are handed over.
struct InputFormat {
    enum Kind : int { RGB48 = 0, Planar = 1, XyzPlanar = 2 };
    int kind = XyzPlanar;
    int width = 0, height = 0;                     // container
    int aw = 0, ah = 0, ox = 0, oy = 0;            // planar: active picture and its offset
    int sx = 1, sy = 1, bits = 8, rgb = 0, fullRange = 0, chroma = 0;
    double kr = 0.2126, kb = 0.0722;
    size_t bytes() const;
};

struct FrameConfig {
    Layout layout;
    InputFormat input;
    ColorSpec color;
    // rate control on the device: the frame's byte budget and component cap (0: on the host)
    unsigned long long rcMaxTotal = 0, rcMaxComp = 0;
};

struct JobOptions {
    int dwtTile = kDwtDefaultShape;  // tile shape index (see dwt_tile_size): level 0 in groups of 4 120 x 8 strips
    int cmWaves = 4;          // waves per context-modelling workgroup
    int cmGroupsPerCU = 8;    // persistent grid size per compute unit
    int mqGroup = 256;        // MQ workgroup size
    int mqRead = 64;          // symbol bytes each MQ lane reads at a time (16 or 64)
    int mqGroupsPerCU = 8;
    int mqLongPerCU = 8;      // longest blocks per compute unit coded one per wave
    double poolBytesPerCoef = 8.0;  // symbol pool sizing (grown on overflow)
    bool useGraph = false;    // replay the launch chain as a HIP graph
};

// Per-stage device time (ms) of the last run, from HIP events on the job's stream.
struct StageTimes {
    float upload = 0, color = 0, dwt = 0, dwt0 = 0, t1cm = 0, t1mq = 0, compact = 0, download = 0, gpu = 0;
    double symbols = 0, bytes = 0;
    int reruns = 0;
};

// One frame's Tier-1 result, in the job's pinned host buffers (valid until the next run).
struct Tier1View {
    const uint32_t* info5;  // per block: nbps, passes, data offset, data length, pass base
    const uint32_t* passLen;
    const int32_t* passDist;
    const uint8_t* data;
    size_t numBlocks, numPasses, numData;
};

// Device-wide state shared by every job on a device: the constant Tier-1 tables and the
// colour / entropy lookup tables.
// A layout's packet tables on the device (built once per layout and device).
struct DevRcTables {
    const RcPacket* packets = nullptr;
    const RcPBand* pbands = nullptr;
    uint32_t npackets = 0, ntp = 0;
    int tpComp[6] = {};
    unsigned long long fixedBytes = 0;
};

class DeviceContext {
public:
    static DeviceContext& get(int device);
    int device() const { return dev_; }
    const float* log2Table() const { return log2_; }
    const float* thresholds() const { return thr_; }
    const unsigned long long* octaveMantissa() const { return octM_; }
    const uint16_t* octaveFirst() const { return octF_; }
    const DevRcTables& rcTables(const Layout& L);
    int computeUnits() const { return cus_; }
    std::string name() const { return name_; }
    std::string arch() const { return arch_; }

private:
    explicit DeviceContext(int device);
    int dev_ = 0, cus_ = 0;
    float* log2_ = nullptr;
    float* thr_ = nullptr;
    unsigned long long* octM_ = nullptr;
    uint16_t* octF_ = nullptr;
    std::mutex rcMutex_;
    std::map<std::string, DevRcTables> rc_;
    std::string name_, arch_;
};

// Stage times (ms) the autotuner measured for each choice: the DWT tile shapes of
// dwt_tile_size (64x64, 64x32, 32x64, 64x16 LDS tiles, 120x64, 120x32, 120x128 strips,
// 120x16 and 120x8 preloaded strips, the 120x8 strips addressed as buffers, groups of four
// of them writing whole lines, all at level 0);
// context-modelling waves per workgroup 1, 2, 4, 8; MQ lanes per workgroup 64, 128, 256 with
// 16-byte symbol reads, then the same with 64-byte reads.
struct TuneTimes {
    double dwt[kDwtShapes] = {}, cm[4] = {}, mq[6] = {};  // mq: 16-byte reads x 64/128/256 lanes, then 64-byte
};

// Where a W x H frame's wavelet coefficients lie in its device planes: Mallat order (the LL
// quadrant of level 0 holds the deeper levels at their Mallat positions), but every row
// starts on a 128-byte line and level 0's high bands (HL, HH) start at a line too, so the
// level-0 transform writes whole lines. Widths in floats.
struct CoefLayout {
    size_t hiOff;     // column of level 0's HL / HH: ceil(W / 2) rounded up to 32
    size_t stride;    // row stride: hiOff + floor(W / 2) rounded up to 32
    size_t llStride;  // row stride of level 0's LL output: ceil(W / 2) rounded up to 32
    static CoefLayout of(size_t W) {
        auto up32 = [](size_t v) { return (v + 31) & ~(size_t)31; };
        CoefLayout c;
        c.hiOff = up32((W + 1) / 2);
        c.stride = up32(c.hiOff + W / 2);
        c.llStride = c.hiOff;
        return c;
    }
};

// Room one frame slot of a job holds: any frame up to maxWidth x maxHeight, and (colour)
// the code planes and EOTF table that planar or RGB48 inputs need.
struct SlotCapacity {
    int maxWidth = 0, maxHeight = 0;
    bool colour = false;
    bool fits(const FrameConfig& f) const;
};

class Job {
public:
    // A job with slots of the given capacities and nothing bound yet.
    Job(int device, std::vector<SlotCapacity> slots, const JobOptions& opt);
    // A job whose slots are exactly these frames, bound.
    Job(int device, std::vector<FrameConfig> frames, const JobOptions& opt);
    ~Job();
    Job(const Job&) = delete;
    Job& operator=(const Job&) = delete;

    // Assigns frames to slots 0..frames.size()-1 for the following runs (frame i must fit
    // slot i). Waits for the job's previous run; the tables go to the device on the job's
    // stream ahead of the next launch.
    void bind(std::vector<FrameConfig> frames);

    int numSlots() const { return (int)slots_.size(); }
    const SlotCapacity& slot(int i) const { return slots_[i]; }
    int numFrames() const { return (int)frames_.size(); }
    const FrameConfig& frame(int i) const { return frames_[i]; }
    hipStream_t stream() const { return stream_; }
    // Device memory this job holds.
    size_t deviceBytes() const { return bytes_; }
    // Device memory a job with these slots (or frames) would take before any growth.
    static size_t estimateBytes(const std::vector<SlotCapacity>& slots, const JobOptions& opt);
    static size_t estimateBytes(const std::vector<FrameConfig>& frames, const JobOptions& opt);

    // Pinned host staging buffer of slot i (room for any input the slot takes), allocated
    // on first use.
    void* hostInput(int i);
    // Copies the bound frame i from host memory into its device input buffer (synchronously).
    void upload(int i, const void* frame);
    // The same on the job's stream from the slot's pinned staging buffer, ahead of the
    // next launch (the staging buffer must hold the frame until that run is done).
    void uploadAsync(int i);
    // The same from any host or device buffer (kept valid by the caller until the run is
    // done); a device buffer is copied within HBM.
    void uploadFrom(int i, const void* src);
    // Waits on the host until the copies queued by uploadFrom / uploadAsync have arrived
    // (the job's stream holds nothing else before a launch).
    void inputsReady();
    // Early termination threshold of frame i (0: code every pass).
    void setSlope(int i, double t) { slope_[i] = (float)t; }

    // Runs the whole chain on the bound frames. uploadInputs: copy every frame's pinned
    // staging buffer to the device first (inside the timed chain); otherwise the frames
    // already resident in HBM are encoded. Blocks until the results are in the job's pinned
    // host buffers.
    void encode(bool uploadInputs, StageTimes* times = nullptr);
    // The same in two halves, so that host work can overlap the device: launch() enqueues
    // the chain, finish() waits for it and fetches the results.
    void launch(bool uploadInputs);
    void finish(StageTimes* times = nullptr);
    // True once the last launched chain has completed on the device (finish() still has to
    // be called to fetch its results).
    bool done() const;

    Tier1View result(int i) const;
    // Device rate control result of frame i of the last run (nullptr when it had none).
    const RcFrame* rateResult(int i) const;
    // Stage times of the last run (ms) and its symbol count.
    const StageTimes& lastTimes() const { return last_; }
    // Times every DWT tile shape, context-modelling workgroup size and MQ workgroup size on
    // the bound frames (their inputs must be on the device), keeps the fastest of each for
    // the following runs and returns them. Each choice runs reps times after a warm-up.
    JobOptions autotune(int reps = 2, TuneTimes* times = nullptr);
    const JobOptions& options() const { return opt_; }

    // Debug taps (synchronous): X'Y'Z' codes (3 planes), the DWT output (3 Mallat planes),
    // and the quantised coefficient words (1,024 per block) of frame i.
    void debugCodes(int i, std::vector<uint16_t>& out);
    void debugWavelet(int i, std::vector<float>& out);
    void debugCoeffs(int i, std::vector<uint32_t>& out);
    // The raw per-block symbol counts and pass tables of frame i (for tests).
    void debugInfo(int i, std::vector<T1Info>& out);

private:
    struct SlotDev {
        size_t inCap = 0, pixCap = 0;
        void* in = nullptr;          // device input
        uint16_t* codesBuf = nullptr;  // 3 planes of codes for colour inputs
        float* coef = nullptr;       // 3 coefficient planes (CoefLayout)
        float* llA = nullptr;        // LL scratch, 3 x CoefLayout::llStride x ceil(H/2)
        float* llB = nullptr;        // LL scratch, 3 x ceil(W/4) x ceil(H/4)
        float* lut = nullptr;        // EOTF table (colour slots only)
        float* hostLut = nullptr;    // pinned copy of the table last uploaded
        bool lutValid = false;
        ColorSpec lutSpec;
        // bound frame
        size_t inBytes = 0, npix = 0;
        size_t coefStride = 0, coefPlane = 0, hiOff = 0, llStride = 0;  // CoefLayout of the frame
        uint16_t* codes = nullptr;   // == in for X'Y'Z' planar input
        float params[kColorParams] = {};  // colour parameter block of the bound spec
        float* lut3 = nullptr;             // app 3D LUT on the device (colour slots, on demand)
        size_t lut3Cap = 0;                // its capacity in floats
        uint32_t firstBlock = 0, numBlocks = 0;
    };

    void allocate();
    void release();
    void growT1(bool pool, bool out);
    void enqueueChain(bool uploadInputs);
    void planLevels();
    void enqueueDwt(hipStream_t s);
    void enqueueT1(hipStream_t s);
    template <class T> T* devAlloc(size_t count);

    int dev_;
    DeviceContext& ctx_;
    JobOptions opt_;
    std::vector<SlotCapacity> slots_;
    std::vector<SlotDev> sd_;
    std::vector<FrameConfig> frames_;
    std::vector<void*> hostIn_;
    std::vector<float> slope_;
    std::vector<void*> allocs_;
    size_t bytes_ = 0;
    hipStream_t stream_ = nullptr;
    // uploadFrom puts odd slots' copies on a second stream: two copy queues keep the link busy
    // while the other turns around between copies; the next launch waits for evUp_
    hipStream_t upStream_ = nullptr;
    hipEvent_t evUp_ = nullptr;
    bool upPending_ = false;
    hipEvent_t ev_[9] = {};  // ev_[8]: after DWT level 0
    hipEvent_t evDone_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    bool graphUpload_ = false;
    bool capturing_ = false;
    bool launched_ = false;
    StageTimes last_;

    // capacities
    uint32_t blockCap_ = 0, bandCap_ = 0;
    size_t coefCap_ = 0;
    int levelCap_ = 0;
    // bound composition
    uint32_t nblocks_ = 0;
    int numLevels_ = 0;
    std::vector<int> levelCount_, levelTiles_, levelShape_;
    // job-wide tables (device, and pinned staging for their upload)
    T1Block* dBlocks_ = nullptr;
    T1Block* hBlocks_ = nullptr;
    T1Band* dBands_ = nullptr;
    T1Band* hBands_ = nullptr;
    T1FrameRef* dFrames_ = nullptr;
    T1FrameRef* hFrames_ = nullptr;  // copied before each run (it carries the thresholds)
    DwtLevelDesc* dLevels_ = nullptr;  // levelCap_ x slots
    DwtLevelDesc* hLevels_ = nullptr;
    // device rate control
    bool anyRc_ = false;
    double* dBandScale_ = nullptr;
    double* hBandScale_ = nullptr;
    RcFrame* dRc_ = nullptr;
    RcFrame* hRc_ = nullptr;         // parameters staged by bind(), results after finish()
    double* dHullSlope_ = nullptr;   // hull points, point-major
    double* dHullGain_ = nullptr;
    uint32_t* dHullSel_ = nullptr;
    int32_t* dHullKbin_ = nullptr;
    uint32_t* dHullN_ = nullptr;
    uint32_t* dRcSel_ = nullptr;
    unsigned long long* dRcPart_ = nullptr;
    uint32_t* dRcSegsum_ = nullptr;
    uint32_t rcMaxBlocks_ = 0;       // most blocks of a rate-controlled frame
    uint32_t* dRcHist_ = nullptr;
    // Tier-1 arrays
    T1Info* dInfo_ = nullptr;
    uint32_t* dPassEnd_ = nullptr;
    int32_t* dPassDist_ = nullptr;
    uint32_t* dPassLen_ = nullptr;
    uint2* dChunks_ = nullptr;
    MqSnapshot* dSnaps_ = nullptr;
    uint32_t* dPerm_ = nullptr;
    uint32_t* dHist_ = nullptr;
    unsigned long long* dOutOff_ = nullptr;
    unsigned long long* dDataOff_ = nullptr;
    unsigned long long* dPartial_ = nullptr;
    uint32_t* dCodeLen_ = nullptr;
    JobCounters* dCounters_ = nullptr;
    uint32_t* dWork_ = nullptr;
    uint8_t* dPool_ = nullptr;
    size_t poolCap_ = 0;
    uint8_t* dOut_ = nullptr;
    size_t outCap_ = 0;
    // dense results alias the pool once MQ has consumed it
    uint32_t* dDenseInfo_ = nullptr;
    uint32_t* dDensePassLen_ = nullptr;
    int32_t* dDensePassDist_ = nullptr;
    uint8_t* dDenseData_ = nullptr;
    // pinned host results
    uint32_t* hInfo_ = nullptr;
    uint32_t* hPassLen_ = nullptr;
    int32_t* hPassDist_ = nullptr;
    uint8_t* hData_ = nullptr;
    size_t hDataCap_ = 0, hPassCap_ = 0;
    unsigned long long* hSmall_ = nullptr;  // counters + totals
    size_t numPassesTotal_ = 0, numDataTotal_ = 0;
    std::vector<size_t> frameDataBase_, framePassBase_;
};

// Throws with the HIP error text.
void hipCheck(hipError_t e, const char* what);

}  // namespace gail_amd
