/* The interface between libgail and its HIP (MI355X / gfx950) backend plugin.
 *
 * Drop-in counterpart of the reference's CUDA plugin table (src/cuda/cuda_plugin.h:144-195,
 * entry gail_cuda_plugin_get at cuda_plugin.h:195): libgail dlopens libgail_hip.so from its
 * own directory, calls gail_hip_plugin_get(GAIL_HIP_PLUGIN_ABI) and reaches the backend
 * only through the table below. Only plain data crosses: no C++ types, no exceptions, no
 * torch. A call that can fail returns a GailHipStatus and leaves its message in
 * last_error() for the calling thread. Memory the plugin hands out (names, result views)
 * stays valid until the next call on the same object or until it is destroyed.
 *
 * One GPU only: device_count() reports the devices HIP sees, and the library uses one.
 */
#ifndef GAIL_HIP_PLUGIN_H
#define GAIL_HIP_PLUGIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GAIL_HIP_PLUGIN_ABI 2u
#define GAIL_HIP_PLUGIN_ENTRY "gail_hip_plugin_get"

/* GailCudaStatus (cuda_plugin.h:52-63). */
enum GailHipStatus {
    GAIL_HIP_OK = 0,
    GAIL_HIP_ERROR = 1,         /* std::runtime_error on the library's side */
    GAIL_HIP_DEVICE_MEMORY = 2, /* gail::DeviceOutOfMemory: the device stays usable */
    GAIL_HIP_HOST_MEMORY = 3,   /* std::bad_alloc */
    GAIL_HIP_MISUSE = 4,        /* std::logic_error */
    GAIL_HIP_NO_CODE = 5,       /* device_usable(): no code object for this GPU */
    GAIL_HIP_UNSUPPORTED = 6    /* a feature this backend does not have */
};

typedef void (*GailHipLog)(void* user, int level, const char* message);

typedef struct GailHipDevice GailHipDevice;
typedef struct GailHipEncoder GailHipEncoder;
typedef struct GailHipBatch GailHipBatch;
typedef struct GailHipDecoder GailHipDecoder;

/* GailCudaDeviceDesc (cuda_plugin.h:72-79); gfx arch instead of compute capability. */
typedef struct GailHipDeviceDesc {
    int32_t index;
    int32_t computeUnits;
    uint64_t memoryBytes;
    char name[256];
    char arch[64];      /* "gfx950" */
    char pciBusId[32];  /* "0000:01:00.0" */
} GailHipDeviceDesc;

/* gail::Subband (layout.h:29-40), member for member. */
typedef struct GailHipSubband {
    int32_t comp, res, level, orient;
    int32_t px0, py0, w, h;
    int32_t ncbx, ncby, cblkBase;
    int32_t eps, mu, Mb;
    float delta;
    double weight;
} GailHipSubband;

/* GailCudaLayout (cuda_plugin.h:83-88). profile: 0 CINEMA2K, 1 CINEMA4K. */
typedef struct GailHipLayout {
    int32_t profile, width, height, levels, guardBits;
    int32_t numBands, numCblks;
    double baseStep;
    const GailHipSubband* bands;
} GailHipLayout;

/* gail::FrameFormat (color.h:47-73) plus the planar 12-bit X'Y'Z' layout this backend
 * takes directly. kind: 0 RGB48 (interleaved 16-bit R'G'B'), 1 planar Y'CbCr / G'B'R',
 * 2 planar X'Y'Z' (three planes of 16-bit words holding 12-bit codes). chroma: 0 left,
 * 1 center, 2 top-left. */
typedef struct GailHipFrameFormat {
    int32_t kind, chroma;
    int32_t width, height;
    int32_t activeW, activeH, ox, oy;
    int32_t sx, sy, rgb, bits, fullRange, pad;
    double kr, kb;
} GailHipFrameFormat;

/* GailCudaColor (cuda_plugin.h:91-97), gail::ColorParams (color.h:84-100) in its
 * enumerations: transfer 0 BT.1886, 1 sRGB, 2 PQ, 3 HLG, 4 linear, 5 gamma 2.6; primaries
 * 0 BT.709, 1 BT.2020, 2 P3-D65, 3 DCI-P3, 4 X'Y'Z'; toneMap (PQ / HLG sources) 0 signal,
 * 1 shoulder, 2 BT.2390, 3 none; toneNorm 0 each channel, 1 max(R,G,B), 2 luminance; gamut
 * 0 P3-D65, 1 DCI-P3, 2 none. An app 3D LUT (lut3dData: lut3dSize^3 x 3 floats, red
 * fastest; or a .cube file named by lut3d) replaces the managed route. Non-positive
 * luminances and knee take ColorParams' defaults. */
typedef struct GailHipColor {
    int32_t transfer, primaries, toneMap, toneNorm, gamut, lut3dSize;
    double sdrWhite, hdrPeak, cinemaPeak, hdrRefWhite, cinemaRefWhite, knee, gamutThreshold;
    const char* lut3d;        /* path of a .cube file, or NULL / "" */
    const float* lut3dData;
    uint64_t lut3dCount;      /* floats at lut3dData */
} GailHipColor;

/* GailCudaEncodeSettings (cuda_plugin.h:100-105). */
typedef struct GailHipEncodeSettings {
    int32_t profile, width, height;
    double baseStep, slopeFactor, bitsPerSymbol;
    GailHipFrameFormat input;
    GailHipColor color;
    /* Rate control on the device (no counterpart in the CUDA plugin): the frame's byte budget
     * and per-component cap (rate_control.cpp:11-26); 0 leaves rate control to the host. The
     * frame's Tier-1 output then holds only the chosen passes, and GailHipRate its result. */
    uint64_t rcMaxTotal, rcMaxComp;
} GailHipEncodeSettings;

/* GailCudaMemory (cuda_plugin.h:113-116). */
typedef struct GailHipMemory {
    uint64_t total, available;
    int32_t budgeted;
} GailHipMemory;

/* gail::CblkInfo (tier1_types.h:12-18). */
typedef struct GailHipCblkInfo {
    uint32_t nbps, numPasses, dataOffset, dataLen, passBase;
} GailHipCblkInfo;

/* GailCudaTier1 (cuda_plugin.h:119-125): a Tier1Frame as the encoder left it. */
typedef struct GailHipTier1 {
    const GailHipCblkInfo* info;
    const uint32_t* passLen;
    const int32_t* passDist;
    const uint8_t* data;
    uint64_t numInfo, numPasses, numData;
} GailHipTier1;

/* Result of the device rate control of one frame (no counterpart in the CUDA plugin). */
typedef struct GailHipRate {
    double lambda[3];          /* slope threshold per component */
    uint64_t total, comp[3], tilePart[6];  /* codestream bytes, as Tier-2 will write them */
    double distortion;         /* weighted squared error left (rate-control units) */
    uint32_t iterations, numTileParts;
} GailHipRate;

/* gail::FrameTimings (backend.h:33-37), milliseconds of device time per stage. */
typedef struct GailHipFrameTimings {
    double upload, color, dwt, quant, t1, t1cm, t1mq, download, gpu, total;
    double symbols, codedBytes;
} GailHipFrameTimings;

/* GailCudaDebug (cuda_plugin.h:128-134). full: also coefficients and wavelet planes.
 * xyz12 here is planar (3 planes of codes). */
typedef struct GailHipDebug {
    int32_t full;
    const uint16_t* xyz12;
    const uint32_t* coeffs;
    const float* wavelet;
    uint64_t numXyz12, numCoeffs, numWavelet;
} GailHipDebug;

/* GailCudaDecodeSettings (cuda_plugin.h:108-110), gail::DecodeSettings (decode_backend.h:20-23):
 * reduce = resolution levels left out (the picture is 1/2^reduce of the size); highBits != 0
 * puts each 12-bit code in the high bits of its 16-bit word, else in the low bits. */
typedef struct GailHipDecodeSettings {
    int32_t reduce, highBits;
} GailHipDecodeSettings;

/* gail::CblkCode (t2_parse.h:17-22), 12 bytes: a code-block's bytes within the codestream,
 * its coding passes (0: nothing coded) and magnitude bit-planes. */
typedef struct GailHipCblkCode {
    uint32_t offset, len;
    uint8_t passes, nbps, pad[2];
} GailHipCblkCode;

/* GailCudaDecodeTimings (cuda_plugin.h:137-142): milliseconds of device time for codestream
 * upload + Tier-1 + dequantisation, the inverse DWT, the inverse ICT + picture download, and
 * the call's wall time. wantWavelet is set by the caller; wavelet[c] then points to the
 * dequantised planes before the inverse DWT (numWavelet samples each), valid until the next
 * call on the decoder. */
typedef struct GailHipDecodeTimings {
    double t1, dwt, color, total;
    int32_t wantWavelet;
    const float* wavelet[3];
    uint64_t numWavelet;
} GailHipDecodeTimings;

/* Tuning of a batched job (no counterpart in the CUDA plugin). Zero fields take defaults. */
typedef struct GailHipBatchOptions {
    int32_t dwtTile;      /* 0 64x64, 1 64x32, 2 32x64, 3 64x16 (LDS tiles); 4 120x64, 5 120x32,
                             6 120x128 (register-streaming strips, one wave each); 7, 8 level 0
                             as strips with their input rows loaded ahead (7 120x16, 8 120x8),
                             9 the 120x8 strips addressed as buffers, 10 groups of four such
                             strips writing whole 128-byte lines through LDS (the default);
                             the other levels as 64x16 LDS tiles */
    int32_t cmWaves;      /* waves per context-modelling workgroup: 1, 2, 4, 8 */
    int32_t mqGroup;      /* MQ workgroup size: 64, 128, 256 */
    int32_t useGraph;     /* replay the launch chain as a HIP graph */
    int32_t mqRead;       /* symbol bytes an MQ lane reads at a time: 16 or 64 */
    int32_t reserved;
    double poolBytesPerCoef;
} GailHipBatchOptions;

/* Room of one frame slot of a batch made by slots_create (no counterpart in the CUDA
 * plugin): frames up to maxWidth x maxHeight; colour != 0 also takes planar / RGB48 inputs. */
typedef struct GailHipSlot {
    int32_t maxWidth, maxHeight, colour, reserved;
} GailHipSlot;

struct GailHipPlugin {
    uint32_t abi;      /* GAIL_HIP_PLUGIN_ABI the plugin was built with */
    uint32_t size;     /* sizeof(GailHipPlugin) */
    uint64_t layout;   /* gail_hip_plugin_layout() of the plugin's build */
    const char* name;     /* "gail_hip" */
    const char* toolkit;  /* HIP runtime version it was built with */
    const char* archs;    /* "gfx950" */

    const char* (*last_error)(void);

    /* devices (cuda_plugin.h:151-173) */
    int (*device_count)(void);
    int (*device_desc)(int index, GailHipDeviceDesc* out);
    int (*device_usable)(int index);
    int (*device_create)(int index, GailHipLog log, void* user, GailHipDevice** out);
    void (*device_destroy)(GailHipDevice* d);
    const char* (*device_name)(GailHipDevice* d);
    int (*device_memory)(GailHipDevice* d, GailHipMemory* out);

    /* encoder slots (cuda_plugin.h:175-183): one frame at a time */
    uint64_t (*encoder_bytes)(GailHipDevice* d, const GailHipLayout* layout, const GailHipEncodeSettings* settings);
    int (*encoder_create)(GailHipDevice* d, const GailHipLayout* layout, const GailHipEncodeSettings* settings,
                          GailHipEncoder** out);
    void (*encoder_destroy)(GailHipEncoder* e);
    void* (*encoder_input)(GailHipEncoder* e);
    void (*encoder_set_slope)(GailHipEncoder* e, double threshold);
    /* frame: host memory of settings.input.bytes(). timings and debug may be NULL. */
    int (*encoder_encode)(GailHipEncoder* e, const void* frame, GailHipTier1* out, GailHipFrameTimings* timings,
                          GailHipDebug* debug);

    /* decoder slots (cuda_plugin.h:183-190, the slot of decode_backend.h:42-51): one frame at
     * a time per decoder; decoders on other threads run concurrently. decoder_decode takes
     * the frame's layout (the decoder's geometry with the frame's own step sizes), the
     * codestream (host or device memory) and one GailHipCblkCode per code-block of the
     * layout; blocks with no passes or bytes past `size` decode as zero. dst receives the
     * decoded picture (the layout's size halved `reduce` times, rounding up) as three 16-bit
     * words per pixel (X', Y', Z'), rows strideBytes apart, in host or device memory. The
     * call returns when the picture is there. timings may be NULL. */
    uint64_t (*decoder_bytes)(GailHipDevice* d, const GailHipLayout* layout, const GailHipDecodeSettings* settings);
    int (*decoder_create)(GailHipDevice* d, const GailHipLayout* layout, const GailHipDecodeSettings* settings,
                          GailHipDecoder** out);
    void (*decoder_destroy)(GailHipDecoder* dec);
    int (*decoder_decode)(GailHipDecoder* dec, const GailHipLayout* frame, const uint8_t* data, size_t size,
                          const GailHipCblkCode* blocks, size_t numBlocks, void* dst, ptrdiff_t strideBytes,
                          GailHipDecodeTimings* timings);

    /* ---- batched form (no counterpart in the CUDA plugin) -------------------------------
     * A batch is a fixed set of frame slots of any mix of geometries and input formats,
     * encoded together: one DWT launch per level across all frames and one persistent
     * Tier-1 launch across all their code-blocks. */
    uint64_t (*batch_bytes)(GailHipDevice* d, const GailHipLayout* layouts, const GailHipEncodeSettings* settings,
                            int count, const GailHipBatchOptions* options);
    int (*batch_create)(GailHipDevice* d, const GailHipLayout* layouts, const GailHipEncodeSettings* settings,
                        int count, const GailHipBatchOptions* options, GailHipBatch** out);
    void (*batch_destroy)(GailHipBatch* b);
    /* Pinned staging buffer of frame slot i. */
    void* (*batch_input)(GailHipBatch* b, int i);
    /* Copies frame slot i from host memory into HBM now (outside any timed run). */
    int (*batch_upload)(GailHipBatch* b, int i, const void* frame);
    void (*batch_set_slope)(GailHipBatch* b, int i, double threshold);
    /* Encodes every slot. from_staging: copy each slot's staging buffer to HBM first;
     * otherwise the frames already resident in HBM are encoded. outs[count]; timings (whole
     * batch) may be NULL. */
    int (*batch_encode)(GailHipBatch* b, int from_staging, GailHipTier1* outs, GailHipFrameTimings* timings);
    /* The same split in two: launch returns once the work is queued on the batch's
     * stream; wait blocks until the results are in host memory. */
    int (*batch_launch)(GailHipBatch* b, int from_staging);
    int (*batch_wait)(GailHipBatch* b, GailHipTier1* outs, GailHipFrameTimings* timings);
    /* Debug taps of slot i after a run: planar codes, wavelet planes, coefficient words. */
    int (*batch_debug)(GailHipBatch* b, int i, GailHipDebug* debug);
    /* Device time (ms, HIP events on the batch's stream) of the last run's stages:
     * [0] upload, [1] colour, [2] DWT (all levels), [3] DWT level 0, [4] context modelling,
     * [5] ordering + MQ, [6] compaction, [7] download, [8] whole run; [9] Tier-1 symbols,
     * [10] codeword bytes, [11] re-runs after a buffer overflow. Fills up to n values. */
    int (*batch_stage_ms)(GailHipBatch* b, double* out, int n);

    /* Batches whose frame slots are bound per run (the served form). slots_create makes
     * count empty slots; batch_bind assigns frames 0..count-1 (layouts and settings of any
     * mix that fits the slots) for the following runs; it waits for the batch's previous run.
     * batch_input(b, i) is valid for every slot; the other batch entries act on the bound
     * frames. */
    uint64_t (*slots_bytes)(GailHipDevice* d, const GailHipSlot* slots, int count, const GailHipBatchOptions* options);
    int (*slots_create)(GailHipDevice* d, const GailHipSlot* slots, int count, const GailHipBatchOptions* options,
                        GailHipBatch** out);
    int (*batch_bind)(GailHipBatch* b, const GailHipLayout* layouts, const GailHipEncodeSettings* settings, int count);
    /* 1 once the last launched run has finished on the device (batch_wait still fetches
     * its results), 0 while it runs, a negative status on error. */
    int (*batch_query)(GailHipBatch* b);
    /* Queues the copy of bound frame i from src (its input layout, bytes as the frame's
     * settings give) on the batch's stream, ahead of the next batch_launch(b, 0). src must stay
     * valid until that run is done: pinned host memory (host_alloc) is copied by DMA, device
     * memory (device_alloc) within HBM. */
    int (*batch_upload_from)(GailHipBatch* b, int i, const void* src);
    /* Pinned host memory on the plugin's runtime; NULL when none is left. */
    void* (*host_alloc)(uint64_t bytes);
    void (*host_free)(void* p);
    /* Autotuner: times every DWT tile shape (dwtTile 0-10), context-modelling waves per
     * workgroup (1, 2, 4, 8) and MQ lanes per workgroup (64, 128, 256) on the bound frames,
     * whose inputs must already be on the device; keeps the fastest of each for the batch's
     * later runs, and the MQ symbol read width (16 or 64 bytes). chosen (may be NULL)
     * receives them; times (may be NULL) up to 21 measured stage times in ms: tile shapes
     * 0-3, 4 workgroup sizes, 3 MQ workgroup sizes with 16-byte reads, the same with 64-byte,
     * then tile shapes 4-10. */
    int (*batch_autotune)(GailHipBatch* b, int reps, GailHipBatchOptions* chosen, double* times, int ntimes);
    /* Device memory for frames kept resident in HBM (a frame whose planes lie there is
     * copied device to device into its slot), a synchronous copy in any direction, and
     * whether an address is device memory (1), host memory (0). */
    void* (*device_alloc)(GailHipDevice* d, uint64_t bytes);
    void (*device_free)(void* p);
    int (*copy)(void* dst, const void* src, uint64_t bytes);
    int (*is_device_pointer)(const void* p);
    /* Device rate control result of frame i of the last run (GAIL_HIP_MISUSE when the frame's
     * settings left rate control to the host). */
    int (*batch_rate)(GailHipBatch* b, int i, GailHipRate* out);
    /* Blocks until every frame queued by batch_upload_from since the last run has arrived in
     * the batch's slots. A caller that launches only then keeps the device's queues free of
     * waits on copies, and can hand the copy engines one batch's frames at a time. */
    int (*batch_inputs_ready)(GailHipBatch* b);
    /* Device rate control result of the encoder's last frame (GAIL_HIP_MISUSE when its
     * settings left rate control to the host); as batch_rate, for encoder_encode. */
    int (*encoder_rate)(GailHipEncoder* e, GailHipRate* out);
};
typedef struct GailHipPlugin GailHipPlugin;

/* The one symbol the plugin exports (replaces gail_cuda_plugin_get, cuda_plugin.h:195).
 * Returns the table, or NULL for another GAIL_HIP_PLUGIN_ABI. */
__attribute__((visibility("default"))) const GailHipPlugin* gail_hip_plugin_get(uint32_t abi);

/* decoder_decode without a decoder object and without device memory of its own. It has no
 * counterpart in the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its intermediate buffers are
 * a fixed 80 MiB array in the plugin's device code, which each GPU the plugin runs on holds
 * until the process ends; so it takes frames up to 2048 x 1080 (any 2K DCI frame) with at
 * most 1,572,864 bytes of codestream, and refuses larger ones with GAIL_HIP_MISUSE. frame,
 * reduce, highBits, data, size, blocks and strideBytes mean what they mean for
 * decoder_decode; data may be host or device memory, dst must be device memory. stream: a
 * hipStream_t to work on (NULL: the default stream). Synchronous: returns when the picture is
 * in dst. Calls on one GPU share the array (it keeps the last call's contents, and the words
 * it must start from zero are cleared by a kernel on every call); the plugin runs them one at
 * a time. The first call on a device also loads the decoder's constant tables there. */
__attribute__((visibility("default"))) int gail_hip_decode_noalloc(int device, const GailHipLayout* frame, int reduce,
                                                                   int highBits, const uint8_t* data, size_t size,
                                                                   const GailHipCblkCode* blocks, size_t numBlocks,
                                                                   void* dst, ptrdiff_t strideBytes, void* stream);

/* Sizes of every structure above folded into one number (cuda_plugin.h:208-217), which
 * the loader compares with the table's layout field. */
static inline uint64_t gail_hip_plugin_layout(void) {
    const uint64_t sizes[] = {sizeof(GailHipDeviceDesc), sizeof(GailHipSubband),      sizeof(GailHipLayout),
                              sizeof(GailHipFrameFormat), sizeof(GailHipColor),       sizeof(GailHipEncodeSettings),
                              sizeof(GailHipMemory),      sizeof(GailHipCblkInfo),    sizeof(GailHipTier1),
                              sizeof(GailHipFrameTimings), sizeof(GailHipDebug),      sizeof(GailHipBatchOptions),
                              sizeof(GailHipSlot),        sizeof(GailHipRate),        sizeof(GailHipDecodeSettings),
                              sizeof(GailHipCblkCode),    sizeof(GailHipDecodeTimings), sizeof(GailHipPlugin)};
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; ++i) h = (h ^ sizes[i]) * 1099511628211ull;
    return h;
}

#ifdef __cplusplus
}
#endif
#endif
