/* GAIL encoder library API on the HIP backend (libgail.so of this build).
 *
 * The encoder half of the reference's public header (include/gail/gail.h:1-362), with the
 * same names, structure layouts and enum values, so applications written against the
 * reference compile and run unchanged: pull (send / receive), callback (on_packet) and
 * blocking (gail_encoder_encode) encodes. Additions, as gail.h:82 allows ("enum values and
 * functions may be added"): GAIL_BACKEND_HIP, GAIL_PIX_XYZ12_PLANAR, and the helper entry
 * points at the end. The decoder half (gail.h:385-538) is not part of this build.
 */
#ifndef GAIL_AMD_H
#define GAIL_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define GAIL_API __attribute__((visibility("default")))
#else
#define GAIL_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define GAIL_VERSION_MAJOR 1
#define GAIL_VERSION_MINOR 1
#define GAIL_VERSION_PATCH 0
#define GAIL_VERSION ((GAIL_VERSION_MAJOR << 16) | (GAIL_VERSION_MINOR << 8) | GAIL_VERSION_PATCH)

/* gail.h:85-95 */
typedef enum gail_status {
    GAIL_OK = 0,
    GAIL_EAGAIN = 1,
    GAIL_EOF = 2,
    GAIL_ERROR_INVALID = -1,
    GAIL_ERROR_UNSUPPORTED = -2,
    GAIL_ERROR_NOMEM = -3,
    GAIL_ERROR_FAILED = -4,
    GAIL_ERROR_STATE = -5,
    GAIL_ERROR_CANCELLED = -6
} gail_status;

GAIL_API const char* gail_last_error(void);
GAIL_API const char* gail_status_string(int status);
GAIL_API unsigned gail_version(void);
GAIL_API const char* gail_version_string(void);

/* gail.h:107-112 */
typedef enum gail_profile { GAIL_PROFILE_2K = 0, GAIL_PROFILE_4K = 1 } gail_profile;
typedef enum gail_container { GAIL_CONTAINER_FLAT = 0, GAIL_CONTAINER_SCOPE = 1, GAIL_CONTAINER_FULL = 2 } gail_container;
GAIL_API int gail_container_size(gail_profile profile, gail_container container, int* width, int* height);

/* gail.h:124-142; GAIL_PIX_XYZ12_PLANAR: three planes (data[0..2] = X', Y', Z') of 16-bit
 * words holding 12-bit codes in the low bits, encoded exactly. */
typedef enum gail_pixel_format {
    GAIL_PIX_RGB48 = 0,
    GAIL_PIX_XYZ12 = 1,
    GAIL_PIX_XYZ12_MSB = 2,
    GAIL_PIX_YUV420P = 3,
    GAIL_PIX_YUV420P10 = 4,
    GAIL_PIX_YUV420P12 = 5,
    GAIL_PIX_YUV422P = 6,
    GAIL_PIX_YUV422P10 = 7,
    GAIL_PIX_YUV422P12 = 8,
    GAIL_PIX_YUV444P = 9,
    GAIL_PIX_YUV444P10 = 10,
    GAIL_PIX_YUV444P12 = 11,
    GAIL_PIX_GBRP = 12,
    GAIL_PIX_GBRP10 = 13,
    GAIL_PIX_GBRP12 = 14,
    GAIL_PIX_GBRP16 = 15,
    GAIL_PIX_XYZ12_PLANAR = 16
} gail_pixel_format;

typedef enum gail_matrix { GAIL_MATRIX_BT709 = 0, GAIL_MATRIX_BT2020 = 1, GAIL_MATRIX_BT601 = 2 } gail_matrix;
typedef enum gail_chroma_location {
    GAIL_CHROMA_LEFT = 0,
    GAIL_CHROMA_CENTER = 1,
    GAIL_CHROMA_TOPLEFT = 2
} gail_chroma_location;

/* gail.h:152-161 */
typedef struct gail_input {
    gail_pixel_format format;
    int width, height; /* picture size; 0 = the container. A smaller picture is centred. */
    gail_matrix matrix;
    int full_range;
    gail_chroma_location chroma_location;
    uint64_t reserved[4];
} gail_input;

/* gail.h:165-232 */
typedef enum gail_color_mode { GAIL_COLOR_MANAGED = 0, GAIL_COLOR_LUT3D = 1 } gail_color_mode;
typedef enum gail_transfer {
    GAIL_TRANSFER_BT1886 = 0,
    GAIL_TRANSFER_SRGB = 1,
    GAIL_TRANSFER_PQ = 2,
    GAIL_TRANSFER_HLG = 3,
    GAIL_TRANSFER_LINEAR = 4,
    GAIL_TRANSFER_GAMMA26 = 5
} gail_transfer;
typedef enum gail_primaries {
    GAIL_PRIMARIES_BT709 = 0,
    GAIL_PRIMARIES_BT2020 = 1,
    GAIL_PRIMARIES_P3D65 = 2,
    GAIL_PRIMARIES_DCIP3 = 3
} gail_primaries;
typedef enum gail_tone_map {
    GAIL_TONE_MAP_SIGNAL = 0,
    GAIL_TONE_MAP_SHOULDER = 1,
    GAIL_TONE_MAP_BT2390 = 2,
    GAIL_TONE_MAP_NONE = 3
} gail_tone_map;
typedef enum gail_tone_norm {
    GAIL_TONE_NORM_CHANNEL = 0,
    GAIL_TONE_NORM_MAXRGB = 1,
    GAIL_TONE_NORM_LUMINANCE = 2
} gail_tone_norm;
typedef enum gail_gamut { GAIL_GAMUT_NONE = 0, GAIL_GAMUT_P3D65 = 1, GAIL_GAMUT_DCIP3 = 2 } gail_gamut;

typedef struct gail_color {
    gail_color_mode mode;
    gail_transfer transfer;
    gail_primaries primaries;
    gail_tone_map tone_map;
    gail_tone_norm tone_norm;
    double sdr_white;
    double hdr_peak;
    double hdr_ref_white;
    double cinema_ref_white;
    double cinema_peak;
    double knee;
    gail_gamut gamut;
    double gamut_threshold;
    const float* lut3d;
    int lut3d_size;
    uint64_t reserved[8];
} gail_color;

typedef enum gail_preset {
    GAIL_PRESET_SDR_BT709 = 0,
    GAIL_PRESET_SDR_BT2020 = 1,
    GAIL_PRESET_HDR_PQ_BT2020 = 2,
    GAIL_PRESET_HDR_HLG_BT2020 = 3
} gail_preset;
GAIL_API void gail_color_preset(gail_preset preset, gail_color* c);

/* gail.h:245-250, plus the HIP backend */
typedef enum gail_backend {
    GAIL_BACKEND_AUTO = 0,
    GAIL_BACKEND_CUDA = 1,
    GAIL_BACKEND_VULKAN = 2,
    GAIL_BACKEND_CPU = 3,
    GAIL_BACKEND_HIP = 4
} gail_backend;

typedef enum gail_log_level { GAIL_LOG_ERROR = 0, GAIL_LOG_WARNING = 1, GAIL_LOG_INFO = 2, GAIL_LOG_DEBUG = 3 } gail_log_level;

/* gail.h:254-263 */
typedef struct gail_packet {
    uint8_t* data;
    size_t size;
    int64_t index;
    void* opaque;
    int device;
    int compliant;
    double psnr_estimate;
    uint64_t reserved[4];
} gail_packet;
GAIL_API void gail_packet_free(gail_packet* pkt);

/* gail.h:267-272 */
typedef struct gail_frame {
    const void* data[3];
    ptrdiff_t stride[3];
    void* opaque;
    uint64_t reserved[4];
} gail_frame;

typedef void (*gail_packet_cb)(void* user, const gail_packet* pkt);
typedef void (*gail_log_cb)(void* user, gail_log_level level, const char* message);

/* gail.h:277-321 */
typedef struct gail_config {
    size_t struct_size;
    gail_profile profile;
    gail_container container;
    double fps;
    double max_mbps;
    int stereo;
    gail_input input;
    gail_color color;
    gail_backend backend;
    const int* devices;
    int num_devices;
    int frames_in_flight;
    int check;
    gail_packet_cb on_packet;
    void* user;
    gail_log_cb log;
    void* log_user;
    gail_log_level log_level;
    const char* const* device_ids;
    int num_device_ids;
    uint64_t max_device_memory;
    int cpu_threads;
    int low_priority;
} gail_config;

GAIL_API void gail_config_init(gail_config* c, size_t struct_size);
static inline void gail_config_defaults(gail_config* c) { gail_config_init(c, sizeof(gail_config)); }

typedef struct gail_encoder gail_encoder;

/* gail.h:329-362 */
GAIL_API int gail_encoder_create(const gail_config* c, gail_encoder** encoder);
GAIL_API void gail_encoder_destroy(gail_encoder* e);
GAIL_API void gail_encoder_frame_size(const gail_encoder* e, int* width, int* height);
GAIL_API const char* gail_encoder_devices(const gail_encoder* e);
GAIL_API int gail_encoder_send(gail_encoder* e, const gail_frame* frame);
GAIL_API int gail_encoder_receive(gail_encoder* e, gail_packet* pkt, int wait);
GAIL_API int gail_encoder_encode(gail_encoder* e, const gail_frame* frame, gail_packet* pkt);
GAIL_API int gail_encoder_finish(gail_encoder* e);
GAIL_API void gail_encoder_cancel(gail_encoder* e);
GAIL_API const char* gail_encoder_error(const gail_encoder* e);

/* gail.h:364-383 */
typedef struct gail_device_info {
    gail_backend backend;
    int index;
    char name[256];
    char uuid[48];
    char pci_bus_id[32];
    uint64_t memory_bytes;
    uint64_t reserved[8];
} gail_device_info;
GAIL_API int gail_enumerate_devices(gail_backend backend, gail_device_info* devices, int max_devices);

/* ---- served form (no counterpart in gail.h) --------------------------------------------
 * A request loop on one device for frames of many streams, each with its own profile,
 * container, input format and colour. Frames are admitted into device jobs as job slots free
 * up (the number of slots is set by a device-memory budget), the frames admitted together are
 * coded in one job, several jobs run on the device at once, and rate control + Tier-2 run on a
 * host thread pool. Every frame's result is delivered through the callback with its latency. */
typedef struct gail_server gail_server;

typedef struct gail_server_result {
    uint64_t id;                /* the id given to gail_server_submit */
    uint32_t stream;
    int status;                 /* GAIL_OK, or why this frame failed */
    const gail_packet* packet;  /* the codestream; valid during the callback only */
    double queued_ms;           /* submit -> admitted into a device job */
    double latency_ms;          /* submit -> codestream ready */
    int batch_frames;           /* frames in the device job the frame was coded in */
    int recoded;                /* coded again without early termination (lambda < threshold) */
    double lambda;              /* the frame's rate-control slope */
} gail_server_result;

typedef void (*gail_server_cb)(void* user, const gail_server_result* result);

typedef struct gail_server_config {
    size_t struct_size;
    int device;
    int slots_per_job;      /* frame slots per device job (default 8), fewer when one job of
                               a slot class would not fit the memory there is */
    int max_jobs;           /* most device jobs held at once (default 4). Jobs have 2K or 4K
                               slots; they are made as frames of their class wait, while the
                               device has their memory free beyond a headroom of max(512 MiB,
                               6 %), and retired when idle to make room for the other class
                               or when the device runs short of memory */
    uint64_t memory_budget; /* most device bytes the jobs may hold; 0 = no limit of its own */
    int cpu_threads;        /* rate control + Tier-2 + input threads (0 = usable cores - 1) */
    int check;              /* run the DCI codestream check on every frame */
    int batch_window_us;    /* an idle job waits up to this long for more frames (0: launch at once) */
    int host_rate_control;  /* 1: PCRD on the host pool; 0 (default): on the GPU, only the kept
                               bytes of each frame come back and the pool writes Tier-2 */
    gail_server_cb on_result; /* called from pool threads, one call at a time */
    void* user;
} gail_server_config;

typedef struct gail_server_stats {
    uint64_t frames, failed, recoded;
    double seconds;       /* first submit -> last result */
    double fps;
    double p50_ms, p99_ms, max_ms;          /* latency */
    double queued_p50_ms, queued_p99_ms;
    double mean_batch;    /* frames per device job */
    int jobs, slots_per_job, max_frames_in_flight;  /* jobs held now; 4K slots per job */
    uint64_t job_bytes;   /* device memory the jobs hold now */
    uint64_t slot_bytes;  /* device memory of one 4K frame slot */
    int jobs_2k, jobs_4k; /* jobs held now, by slot class */
    int jobs_made, jobs_retired, max_jobs;  /* jobs made, retired, most held at once */
    int slots_per_job_2k;
    uint64_t peak_job_bytes;  /* most device memory the jobs held at once */
    uint64_t slot_bytes_2k;   /* device memory of one 2K frame slot */
} gail_server_stats;

GAIL_API void gail_server_config_init(gail_server_config* c, size_t struct_size);
GAIL_API int gail_server_create(const gail_server_config* c, gail_server** server);
/* Queues one frame of a stream. stream_cfg describes the stream (profile, container, fps,
 * max_mbps, stereo, input, color; the other fields are ignored); frames of one stream id must
 * use the same description. The frame's planes are read when it is admitted: keep them valid
 * until its result arrives. slope: early-termination threshold of this frame, or negative for
 * 0.25 x the stream's last lambda (pipeline.cpp:194-218). Never blocks on the device. */
GAIL_API int gail_server_submit(gail_server* s, uint64_t id, uint32_t stream, const gail_config* stream_cfg,
                                const gail_frame* frame, double slope);
/* Waits until every submitted frame has been delivered. */
GAIL_API int gail_server_drain(gail_server* s);
GAIL_API int gail_server_get_stats(gail_server* s, gail_server_stats* out);
GAIL_API void gail_server_destroy(gail_server* s);
/* Pinned host memory for frames: a frame whose planes lie back to back in the device's input
 * layout is copied to the device without a host copy, by DMA when it is in such memory. */
GAIL_API void* gail_host_alloc(size_t bytes);
GAIL_API void gail_host_free(void* p);
/* Device memory (device 0) for frames kept resident in HBM: a frame whose planes lie there,
 * back to back in the device's input layout, is copied within HBM into its slot. gail_copy
 * copies in any direction (synchronously). */
GAIL_API void* gail_device_alloc(size_t bytes);
GAIL_API void gail_device_free(void* p);
GAIL_API int gail_copy(void* dst, const void* src, size_t bytes);

/* ---- helpers of this build (no counterpart in gail.h) --------------------------------- */
struct GailHipSubband;
/* makeLayout (layout.cpp:181-210) as plain subbands; returns the code-block count. */
GAIL_API int gail_layout_make(int width, int height, int fourk, double base_step, struct GailHipSubband* bands,
                              int cap, int* nbands);
/* The code-blocks of that layout as (band, x0, y0, w, h) rows. */
GAIL_API int gail_layout_cblks(int width, int height, int fourk, int* out5, int cap);
/* Rate control + Tier-2 of one frame's Tier-1 output (pipeline.cpp:210-226). */
GAIL_API int64_t gail_encode_codestream(int width, int height, int fourk, double base_step, const uint32_t* info5,
                                        const uint32_t* pass_len, const int32_t* pass_dist, const uint8_t* data,
                                        uint64_t num_blocks, double fps, double mbps, uint8_t* out, uint64_t cap,
                                        double* lambda, double* mse, const char* tag);
/* Frame sizes from buildCodestream and from the device's packet-size code (run on the host)
 * over `trials` pass choices; the number of choices where they differ, -1 on error. */
GAIL_API int gail_rc_measure_check(int width, int height, int fourk, double base_step, const uint32_t* info5,
                                   const uint32_t* pass_len, const int32_t* pass_dist, const uint8_t* data,
                                   uint64_t num_blocks, int trials, uint64_t seed);
/* Rate control's slope grid: floor(1024 log2 s), and the smallest double of bin k. */
GAIL_API int64_t gail_slope_bin(double s);
GAIL_API double gail_slope_bin_edge(int64_t k);
/* Tier-2 of a Tier-1 frame whose passes were already chosen (each block keeps its coded
 * passes, as the device rate control leaves them); the codestream size, or < 0. */
GAIL_API int64_t gail_build_codestream(int width, int height, int fourk, double base_step, const uint32_t* info5,
                                       const uint32_t* pass_len, const int32_t* pass_dist, const uint8_t* data,
                                       uint64_t num_blocks, uint8_t* out, uint64_t cap, const char* tag);
/* DCI structural check (codestream_check.cpp:43-280); 0 when it passes. */
GAIL_API int gail_check_codestream(const uint8_t* data, uint64_t size, double fps, int fourk, char* msg, int cap);
GAIL_API uint64_t gail_empty_codestream_bytes(int width, int height, int fourk, double base_step);
/* The seeded synthetic test frame: three planes of 12-bit X'Y'Z' codes. */
GAIL_API void gail_synth_xyz12(uint64_t seed, int width, int height, uint16_t* out, int threads);

#ifdef __cplusplus
}
#endif
#endif
