/* CPU oracle for the GAIL encoder hot path: colour conversion, 9/7 DWT, dead-zone
 * quantisation and EBCOT Tier-1 (context modelling, MQ coding, truncation lengths,
 * per-pass distortion, early termination), plus an independent Tier-1 decoder.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (libgail_amd) links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do,
 * as the checker or the timed CPU baseline.
 *
 * Every function restates the algorithm of the reference file:line it cites, written
 * from the description of the algorithm (ISO/IEC 15444-1 Annexes C, D, E, F, G and the
 * reference's behaviour), compiled with -ffp-contract=off so that float results are
 * bit-identical to the reference's x86-64 build (SSE2, no FMA).
 *
 * Parity status per function (how each is pinned; see tests/test_oracle_*.py):
 *   or_layout_make        golden layouts from the compiled reference + closed forms
 *   or_color_*            golden colour samples from the compiled reference; X'Y'Z'
 *                         identity over all 4096 codes
 *   or_fdwt97_2d_f32      golden DWT outputs from the compiled reference; fp64 oracle
 *   or_fdwt97_2d_f64      perfect reconstruction against or_idwt97_2d_f64
 *   or_quantize_block     golden coefficient words + closed form
 *   or_t1_encode          golden Tier-1 results (codeword, pass lengths, distortions,
 *                         early stop) from the compiled reference; round trip through
 *                         or_t1_decode at every truncation point
 *   or_t1_decode          brute-force/round-trip only: independent restatement of
 *                         Annexes C.3 and D (parity pinned by the encoder round trip)
 *   or_t1_decode_words    reconstruction words equal to the compiled reference's
 *                         t1DecodeBlock at every truncation point; closed form
 *                         (2 * magnitude + 2^last plane) against or_t1_decode
 *   or_idwt97_2d_f32      equal to the compiled reference's idwt97_2d; perfect
 *                         reconstruction against or_fdwt97_2d_f32 within float error
 *   or_decode_frame       pictures and wavelet planes equal to the compiled reference's
 *                         CPU decode slot (cpu_decode_backend.cpp)
 */
#ifndef GAIL_ORACLE_H
#define GAIL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_PASSES 64
#define OR_MAX_BANDS (3 * (3 * 6 + 1))

typedef struct or_band {
    int32_t comp, res, level, orient; /* orient: 0 LL, 1 HL, 2 LH, 3 HH */
    int32_t px0, py0, w, h;           /* rectangle in the Mallat-layout plane */
    int32_t ncbx, ncby, cblk_base;
    int32_t eps, mu, mb;
    float delta;
    float pad;
    double weight;
} or_band;

typedef struct or_cblk {
    int32_t band, x0, y0, w, h; /* inside the band */
} or_cblk;

/* layout.cpp:77-210. fourk: 1 = CINEMA4K (NL 6, 2 guard bits), 0 = CINEMA2K (NL 5, 1).
 * Fills up to OR_MAX_BANDS bands and cblk_cap code-blocks; returns the number of
 * code-blocks (even if it exceeds cblk_cap), or -1 on an invalid size. */
int or_layout_make(int width, int height, int fourk, double base_step, or_band* bands, int* nbands,
                   or_cblk* cblks, int cblk_cap);
double or_synthesis_norm(int level, int orient);

/* ---- colour (color.cpp, color_math.h) ---- */
/* transfer: 0 BT1886, 1 sRGB, 2 PQ(signal), 3 HLG(signal), 4 linear, 5 gamma26;
 * primaries: 0 BT709, 1 BT2020, 2 P3D65, 3 DCIP3, 4 XYZ (pass-through).
 * Fills the 65,536-entry EOTF LUT and the 48-float parameter block. */
void or_color_tables(int transfer, int primaries, double sdr_white, double cinema_peak, float* lut65536,
                     float* params48);
/* cm_rgb_to_xyz12 for one R'G'B' pixel given as 16-bit codes; codes out as ints. */
void or_color_pixel(const float* lut, const float* params, uint16_t r, uint16_t g, uint16_t b, int out[3]);
/* ColorParams of the managed route (color.h:84-100) in ColorParams' enumerations:
 * tone_map 0 signal, 1 shoulder, 2 BT.2390, 3 none; tone_norm 0 channel, 1 max(R,G,B),
 * 2 luminance; gamut 0 P3-D65, 1 DCI-P3, 2 none; lut3d_size: nodes per axis of an app 3D
 * LUT (0: none). */
typedef struct or_color_cfg {
    int transfer, primaries, tone_map, tone_norm, gamut, lut3d_size;
    double sdr_white, hdr_peak, cinema_peak, hdr_ref_white, cinema_ref_white, knee, gamut_threshold;
} or_color_cfg;
/* The EOTF LUT and the parameter block for any configuration. */
void or_color_setup(const or_color_cfg* c, float* lut65536, float* params48);
/* cm_rgb_to_xyz12 for n RGB48 pixels (tone map, gamut compression, 3D LUT route). */
void or_color_pixels(const float* lut, const float* params, const float* lut3, int n, const uint16_t* rgb48,
                     int* codes);
/* planarToRgb48Row for a whole container frame (kind planar), into interleaved RGB48.
 * bits: sample depth; sx, sy: chroma subsampling log2; rgb: planes G',B',R';
 * chroma: 0 left, 1 center, 2 top-left; kr, kb: Y'CbCr matrix. */
void or_planar_to_rgb48(const void* src, int bits, int aw, int ah, int sx, int sy, int rgb, int full_range,
                        int chroma, double kr, double kb, int width, int height, int ox, int oy, uint16_t* rgb48);
/* convertFrame: RGB48 -> X'Y'Z' codes (optional) and the three DC-shifted ICT planes. */
void or_convert_frame(const float* lut, const float* params, const uint16_t* rgb48, int width, int height,
                      float* planes /* 3 x W*H */, uint16_t* xyz12 /* may be NULL */);
/* The X'Y'Z' pass-through shortcut: planar 12-bit codes straight to ICT planes. */
void or_xyz_planar_to_ict(const uint16_t* xyz /* 3 planes */, int width, int height, float* planes);
/* fill() of gail_api.cpp:389-394: 12-bit code -> 16-bit signal. */
uint16_t or_xyz12_to_16(uint32_t v);

/* ---- 9/7 DWT (dwt_cpu.cpp) ---- */
void or_fdwt97_2d_f32(float* plane, int width, int height, int levels);
void or_fdwt97_2d_f64(double* plane, int width, int height, int levels);
void or_idwt97_2d_f64(double* plane, int width, int height, int levels);

/* ---- quantisation (quant.h:20-25, quant.cpp:8-20) ---- */
uint32_t or_quantize_coeff(float v, float inv_delta, uint32_t max_mag);
void or_quantize_block(const float* plane, int stride, int x0, int y0, int w, int h, float delta, int mb,
                       uint32_t* out1024);

/* ---- Tier-1 (t1_ref.cpp, mq_encoder.h) ---- */
typedef struct or_t1_result {
    int32_t nbps, num_passes;
    uint32_t len;
    uint32_t pass_len[OR_MAX_PASSES];
    int32_t pass_dist[OR_MAX_PASSES];
} or_t1_result;

/* Codes one 32x32-word block (w x h used). max_passes limits the passes; slope_thresh > 0
 * enables early termination with dist_scale the band's delta^2 * weight. The codeword
 * goes to out (out_cap bytes; 1024 * 4 * 24 + 64 always suffices). Optional symbol trace:
 * syms receives (cx << 1) | d, pass_end the symbol count after each pass. Returns 0, or
 * -1 when the block has more than 22 bit-planes or out_cap is too small. */
int or_t1_encode(const uint32_t* coeffs, int w, int h, int orient, int max_passes, float slope_thresh,
                 float dist_scale, uint8_t* out, uint32_t out_cap, or_t1_result* res, uint8_t* syms,
                 uint32_t sym_cap, uint32_t* nsyms, uint32_t* pass_end);

/* Independent decoder: decodes `passes` coding passes of a codeword of len bytes (the
 * decoder feeds 1-bits past the end). mag receives the decoded magnitude bit-planes
 * (integer part, w*h, row-major), sign the sign bits, sig whether significant. */
int or_t1_decode(const uint8_t* data, uint32_t len, int w, int h, int orient, int nbps, int passes,
                 uint32_t* mag, uint8_t* sign, uint8_t* sig);

/* t1DecodeBlock (t1_dec.cpp:45-181): the reconstruction words of a w x h block (row
 * stride w): bit 31 the sign, the rest twice the decoded magnitude plus the mid-point of
 * the bit-planes not decoded. Returns 0, or -1 for a block size outside 1..32. */
int or_t1_decode_words(const uint8_t* data, uint32_t len, int w, int h, int orient, int nbps, int passes,
                       uint32_t* words);

/* ---- decoder (cpu_decode_backend.cpp, dwt_cpu.cpp:96-135, 213-240) ---- */
/* gail::CblkCode (t2_parse.h:17-22). */
typedef struct or_cblk_code {
    uint32_t offset, len;
    uint8_t passes, nbps, pad[2];
} or_cblk_code;

/* idwt97_2d: levels down to skip + 1 of nplanes float planes. */
void or_idwt97_2d_f32(float* const* planes, int nplanes, int width, int height, int levels, int skip);

/* CpuDecoderSlot::decode (cpu_decode_backend.cpp:38-107) for the layout of (width,
 * height, profile, base_step): Tier-1 and dequantisation of the code-blocks of the
 * resolutions kept, the inverse DWT down to `reduce`, the inverse ICT to 12-bit codes
 * (<< 4 when high_bits), three interleaved words per pixel, rows stride_bytes apart.
 * wavelet (3 x W*H floats, may be NULL) receives the planes before the inverse DWT.
 * Returns 0, or -1 when nblocks or reduce does not fit the layout. */
int or_decode_frame(int width, int height, int fourk, double base_step, const uint8_t* data, size_t size,
                    const or_cblk_code* blocks, int nblocks, int reduce, int high_bits, int threads, uint16_t* dst,
                    ptrdiff_t stride_bytes, float* wavelet);

/* ---- whole frame (cpu_backend.cpp:61-132) ---- */
typedef struct or_cblk_info {
    uint32_t nbps, num_passes, data_offset, data_len, pass_base;
} or_cblk_info;

/* Encodes a frame of three planar 12-bit X'Y'Z' planes (pass-through colour) through
 * ICT, DWT, quantisation and Tier-1 on `threads` threads. Outputs are sized by the
 * caller: info[n], pass_len/pass_dist[n * 64], data[data_cap]; coeffs (n * 1024 words)
 * and planes (3 * W * H floats, the DWT output) may be NULL. slope_thresh as above.
 * Returns the total data bytes, or -1 on overflow. */
int64_t or_encode_frame_xyz(const uint16_t* xyz, int width, int height, int fourk, double base_step,
                            float slope_thresh, int threads, or_cblk_info* info, uint32_t* pass_len,
                            int32_t* pass_dist, uint8_t* data, uint64_t data_cap, uint32_t* coeffs, float* planes);

#ifdef __cplusplus
}
#endif
#endif
