/* CPU oracle for the GAIL encoder hot path. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Built with -O2 -ffp-contract=off so every float expression is evaluated as written,
 * one rounding per operation, exactly like the reference's x86-64 (SSE2) build.
 * Citations are reference file:line; the algorithms are restated, not copied. */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== constants */
/* ISO/IEC 15444-1 Table F.4 lifting constants (dwt97.h:8-12). */
#define L_ALPHA (-1.586134342059924)
#define L_BETA (-0.052980118572961)
#define L_GAMMA 0.882911075530934
#define L_DELTA 0.443506852043971
#define L_K 1.230174104914001
/* ISO/IEC 15444-1 G.3 irreversible component transform (ict.h:11-24). */
static const double ICT_FWD[3][3] = {
    {0.299, 0.587, 0.114}, {-0.16875, -0.33126, 0.5}, {0.5, -0.41869, -0.08131}};
static const double ICT_INV_COLS[3][3] = {{1, 1, 1}, {0, -0.34413, 1.772}, {1.402, -0.71414, 0}};

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ======================================================================== lifting */
/* One lifting step on samples of one parity, whole-sample symmetric extension at both
 * ends (dwt_cpu.cpp:24-33): x[i] += c * (x[i-1] + x[i+1]), with x[-1] := x[1] and
 * x[n] := x[n-2]. n >= 2. */
#define DEFINE_LIFT(NAME, T)                                              \
    static void NAME(T* x, int n, int stride, int parity, T c) {          \
        for (int i = parity; i < n; i += 2) {                             \
            const T l = i > 0 ? x[(size_t)(i - 1) * stride] : x[stride];  \
            const T r = i + 1 < n ? x[(size_t)(i + 1) * stride] : l;      \
            x[(size_t)i * stride] += c * (l + r);                         \
        }                                                                 \
    }
DEFINE_LIFT(lift_f, float)
DEFINE_LIFT(lift_d, double)

/* Forward 1-D transform in place with de-interleaved output (dwt_cpu.cpp:54-93,
 * 146-160): alpha on odd samples, beta on even, gamma odd, delta even, then low-pass
 * (even) samples times 1/K to the front, high-pass (odd) times K behind them. */
#define DEFINE_FWD1D(NAME, T, LIFT)                                                   \
    static void NAME(T* x, int n, int stride, T* tmp) {                              \
        if (n < 2) return;                                                            \
        LIFT(x, n, stride, 1, (T)L_ALPHA);                                            \
        LIFT(x, n, stride, 0, (T)L_BETA);                                             \
        LIFT(x, n, stride, 1, (T)L_GAMMA);                                            \
        LIFT(x, n, stride, 0, (T)L_DELTA);                                            \
        const int sn = (n + 1) / 2;                                                   \
        for (int i = 0; i < n; ++i) {                                                 \
            const T v = x[(size_t)i * stride];                                        \
            if (i & 1) tmp[sn + i / 2] = v * (T)L_K;                                  \
            else tmp[i / 2] = v * (T)(1.0 / L_K);                                     \
        }                                                                             \
        for (int i = 0; i < n; ++i) x[(size_t)i * stride] = tmp[i];                   \
    }
DEFINE_FWD1D(fwd1d_f, float, lift_f)
DEFINE_FWD1D(fwd1d_d, double, lift_d)

/* Inverse 1-D transform (dwt_cpu.cpp:162-178), in place. */
static void inv1d_d(double* x, int n, int stride, double* tmp) {
    if (n < 2) return;
    const int sn = (n + 1) / 2;
    for (int i = 0; i < sn; ++i) tmp[2 * i] = x[(size_t)i * stride] * L_K;
    for (int i = 0; i < n / 2; ++i) tmp[2 * i + 1] = x[(size_t)(sn + i) * stride] * (1.0 / L_K);
    lift_d(tmp, n, 1, 0, -L_DELTA);
    lift_d(tmp, n, 1, 1, -L_GAMMA);
    lift_d(tmp, n, 1, 0, -L_BETA);
    lift_d(tmp, n, 1, 1, -L_ALPHA);
    for (int i = 0; i < n; ++i) x[(size_t)i * stride] = tmp[i];
}

/* Multi-level forward transform in Mallat layout (dwt_cpu.cpp:184-206): per level the
 * columns of the current LL region first, then its rows. */
void or_fdwt97_2d_f32(float* plane, int width, int height, int levels) {
    float* tmp = (float*)malloc(sizeof(float) * (size_t)imax(width, height));
    int w = width, h = height;
    for (int l = 0; l < levels; ++l) {
        for (int x = 0; x < w; ++x) fwd1d_f(plane + x, h, width, tmp);
        for (int y = 0; y < h; ++y) fwd1d_f(plane + (size_t)y * width, w, 1, tmp);
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    free(tmp);
}

void or_fdwt97_2d_f64(double* plane, int width, int height, int levels) {
    double* tmp = (double*)malloc(sizeof(double) * (size_t)imax(width, height));
    int w = width, h = height;
    for (int l = 0; l < levels; ++l) {
        for (int x = 0; x < w; ++x) fwd1d_d(plane + x, h, width, tmp);
        for (int y = 0; y < h; ++y) fwd1d_d(plane + (size_t)y * width, w, 1, tmp);
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    free(tmp);
}

/* Inverse of or_fdwt97_2d_f64: per level rows first, then columns (dwt_cpu.cpp:240-264). */
void or_idwt97_2d_f64(double* plane, int width, int height, int levels) {
    int ws[16], hs[16];
    ws[0] = width;
    hs[0] = height;
    for (int l = 1; l <= levels; ++l) {
        ws[l] = (ws[l - 1] + 1) / 2;
        hs[l] = (hs[l - 1] + 1) / 2;
    }
    double* tmp = (double*)malloc(sizeof(double) * (size_t)imax(width, height));
    for (int l = levels; l >= 1; --l) {
        const int w = ws[l - 1], h = hs[l - 1];
        for (int y = 0; y < h; ++y) inv1d_d(plane + (size_t)y * width, w, 1, tmp);
        for (int x = 0; x < w; ++x) inv1d_d(plane + x, h, width, tmp);
    }
    free(tmp);
}

/* ======================================================================== layout */
/* L2 norm of the 1-D synthesis basis function of a low or high band at `level`
 * (layout.cpp:21-37): a unit impulse in the middle of the band, inverse transformed
 * level by level in double precision. */
static double synth1d(int level, int high) {
    const int n = 64 << level;
    double* sig = (double*)calloc((size_t)n, sizeof(double));
    double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
    const int band_len = n >> level;
    sig[high ? band_len + band_len / 2 : band_len / 2] = 1.0;
    for (int l = level; l >= 1; --l) inv1d_d(sig, n >> (l - 1), 1, tmp);
    double e = 0;
    for (int i = 0; i < n; ++i) e += sig[i] * sig[i];
    free(sig);
    free(tmp);
    return sqrt(e);
}

double or_synthesis_norm(int level, int orient) {
    const int hx = orient == 1 || orient == 3, hy = orient == 2 || orient == 3;
    return synth1d(level, hx) * synth1d(level, hy);
}

/* Synthesis gain of inverse-ICT component c (layout.cpp:53-59). */
static double ict_synthesis_norm(int c) {
    double e = 0;
    for (int i = 0; i < 3; ++i) e += ICT_INV_COLS[c][i] * ICT_INV_COLS[c][i];
    return sqrt(e);
}

/* Geometry (layout.cpp:77-169) and step sizes (layout.cpp:181-210, E.1). */
int or_layout_make(int width, int height, int fourk, double base_step, or_band* bands, int* nbands,
                   or_cblk* cblks, int cblk_cap) {
    if (fourk ? (width > 4096 || height > 2160) : (width > 2048 || height > 1080)) return -1;
    if (width < 64 || height < 64) return -1;
    const int NL = fourk ? 6 : 5, guard = fourk ? 2 : 1;
    int ws[8], hs[8];
    ws[0] = width;
    hs[0] = height;
    for (int l = 1; l <= NL; ++l) {
        ws[l] = (ws[l - 1] + 1) / 2;
        hs[l] = (hs[l - 1] + 1) / 2;
    }
    int nb = 0, count = 0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r <= NL; ++r)
            for (int o = 0; o < (r == 0 ? 1 : 3); ++o) {
                or_band* b = &bands[nb++];
                memset(b, 0, sizeof *b);
                b->comp = c;
                b->res = r;
                b->orient = r == 0 ? 0 : o + 1;
                b->level = r == 0 ? NL : NL - r + 1;
                const int n = b->level;
                const int hx = b->orient == 1 || b->orient == 3, hy = b->orient == 2 || b->orient == 3;
                b->px0 = hx ? ws[n] : 0;
                b->py0 = hy ? hs[n] : 0;
                b->w = hx ? ws[n - 1] - ws[n] : ws[n];
                b->h = hy ? hs[n - 1] - hs[n] : hs[n];
                b->ncbx = (b->w + 31) / 32;
                b->ncby = (b->h + 31) / 32;
                b->cblk_base = count;
                count += b->ncbx * b->ncby;
                /* step size: delta_b ~ base_step / ||synthesis||, signalled as (eps, mu) */
                const int rb = 12 + (b->orient == 0 ? 0 : (b->orient == 3 ? 2 : 1));
                const double norm = or_synthesis_norm(b->level, b->orient);
                const double x = base_step / norm / ldexp(1.0, rb);
                int eps = (int)ceil(-log2(x));
                if (ldexp(x, eps) < 1.0) ++eps;
                if (ldexp(x, eps) >= 2.0) --eps;
                int mu = (int)lround((ldexp(x, eps) - 1.0) * 2048.0);
                if (mu >= 2048) {
                    mu = 0;
                    --eps;
                }
                if (eps < 0 || eps > 31) return -1;
                b->eps = eps;
                b->mu = mu;
                b->delta = (float)(ldexp(1.0, rb - eps) * (1.0 + mu / 2048.0));
                b->mb = guard + eps - 1;
                if (b->mb > 22) return -1;
                const double wc = ict_synthesis_norm(c);
                b->weight = norm * norm * wc * wc;
            }
    *nbands = nb;
    for (int bi = 0; bi < nb; ++bi) {
        const or_band* b = &bands[bi];
        for (int cy = 0; cy < b->ncby; ++cy)
            for (int cx = 0; cx < b->ncbx; ++cx) {
                const int i = b->cblk_base + cy * b->ncbx + cx;
                if (i >= cblk_cap) continue;
                cblks[i].band = bi;
                cblks[i].x0 = cx * 32;
                cblks[i].y0 = cy * 32;
                cblks[i].w = imin(32, b->w - cx * 32);
                cblks[i].h = imin(32, b->h - cy * 32);
            }
    }
    return count;
}

/* ======================================================================== colour */
static void invert3(const double m[9], double r[9]) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    r[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    r[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    r[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    r[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    r[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    r[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    r[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    r[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    r[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

/* Linear RGB (white Y = 1) -> CIE XYZ from the primaries' chromaticities
 * (color.cpp:59-79): columns of P scaled so that RGB = 1 maps to the white point. */
static void rgb_to_xyz_matrix(int primaries, double M[9]) {
    static const double xy[4][4][2] = {
        {{0.640, 0.330}, {0.300, 0.600}, {0.150, 0.060}, {0.3127, 0.3290}}, /* BT.709 */
        {{0.708, 0.292}, {0.170, 0.797}, {0.131, 0.046}, {0.3127, 0.3290}}, /* BT.2020 */
        {{0.680, 0.320}, {0.265, 0.690}, {0.150, 0.060}, {0.3127, 0.3290}}, /* P3-D65 */
        {{0.680, 0.320}, {0.265, 0.690}, {0.150, 0.060}, {0.314, 0.351}},   /* DCI-P3 */
    };
    if (primaries == 4) {
        for (int i = 0; i < 9; ++i) M[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    const double(*c)[2] = xy[primaries];
    double P[9], Pi[9], S[3];
    for (int i = 0; i < 3; ++i) {
        P[i] = c[i][0] / c[i][1];
        P[3 + i] = 1.0;
        P[6 + i] = (1.0 - c[i][0] - c[i][1]) / c[i][1];
    }
    const double W[3] = {c[3][0] / c[3][1], 1.0, (1.0 - c[3][0] - c[3][1]) / c[3][1]};
    invert3(P, Pi);
    for (int i = 0; i < 3; ++i) S[i] = Pi[3 * i] * W[0] + Pi[3 * i + 1] * W[1] + Pi[3 * i + 2] * W[2];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) M[3 * r + k] = P[3 * r + k] * S[k];
}

/* ColorConverter constructor for SDR and pass-through sources (color.cpp:296-372): the
 * EOTF table in cd/m^2 (computed in double, stored as float) and the parameter block
 * (color_params.h:7-23). HDR sources are taken in the default "signal" mode, whose table
 * is the gamma-2.4 one. */
void or_color_tables(int transfer, int primaries, double sdr_white, double cinema_peak, float* lut, float* cp) {
    for (int i = 0; i < 65536; ++i) {
        const double e = i / 65535.0;
        double nits = 0;
        switch (transfer) {
        case 1: nits = (e <= 0.04045 ? e / 12.92 : pow((e + 0.055) / 1.055, 2.4)) * sdr_white; break;
        case 4: nits = e * sdr_white; break;
        case 5: nits = pow(e, 2.6) * 52.37; break;
        default: nits = pow(e, 2.4) * sdr_white; break; /* BT.1886, PQ/HLG as signal */
        }
        lut[i] = (float)nits;
    }
    memset(cp, 0, 48 * sizeof(float));
    double M[9];
    rgb_to_xyz_matrix(primaries, M);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) cp[4 * r + k] = (float)M[3 * r + k];
    cp[17] = (float)cinema_peak;
    cp[42] = 2.f; /* tone norm: channel (default) */
    if (primaries != 4) {
        double T[9], Ti[9];
        rgb_to_xyz_matrix(2, T); /* target gamut P3-D65 (gamut compression off) */
        invert3(T, Ti);
        for (int i = 0; i < 9; ++i) {
            cp[20 + i] = (float)(Ti[i] / cinema_peak);
            cp[29 + i] = (float)(T[i] * cinema_peak);
        }
    }
}

/* cd/m^2 -> 12-bit code (color_math.h:109-112). */
static float encode_code(float v) {
    const float n = fminf(fmaxf(v / 52.37f, 0.0f), 1.0f);
    return floorf(4095.0f * powf(n, 1.0f / 2.6f) + 0.5f);
}

static int lut_index(float e) { return (int)floorf(fminf(fmaxf(e, 0.0f), 1.0f) * 65535.0f + 0.5f); }

/* cm_rgb_to_xyz12, managed route without tone map or gamut compression
 * (color_math.h:117-179; color.cpp:400-407). */
void or_color_pixel(const float* lut, const float* cp, uint16_t r, uint16_t g, uint16_t b, int out[3]) {
    const float R = lut[lut_index(r / 65535.f)];
    const float G = lut[lut_index(g / 65535.f)];
    const float B = lut[lut_index(b / 65535.f)];
    const float X = cp[0] * R + cp[1] * G + cp[2] * B;
    const float Y = cp[4] * R + cp[5] * G + cp[6] * B;
    const float Z = cp[8] * R + cp[9] * G + cp[10] * B;
    out[0] = (int)encode_code(X);
    out[1] = (int)encode_code(Y);
    out[2] = (int)encode_code(Z);
}

/* ---- managed colour with tone map, gamut compression and 3D LUT ---- */
/* SMPTE ST 2084 constants (color.cpp:14-19). */
static const double kPqM1 = 2610.0 / 16384.0, kPqM2 = 2523.0 / 4096.0 * 128.0, kPqC1 = 3424.0 / 4096.0,
                    kPqC2 = 2413.0 / 4096.0 * 32.0, kPqC3 = 2392.0 / 4096.0 * 32.0;

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static double maxd(double a, double b) { return a < b ? b : a; } /* std::max */

/* pqEotf / pqInverseEotf in double (color.cpp:209-218). */
static double pq_eotf_d(double e) {
    e = clampd(e, 0.0, 1.0);
    const double p = pow(e, 1.0 / kPqM2);
    return 10000.0 * pow(maxd(p - kPqC1, 0.0) / (kPqC2 - kPqC3 * p), 1.0 / kPqM1);
}
static double pq_inv_d(double nits) {
    const double y = pow(clampd(nits / 10000.0, 0.0, 1.0), kPqM1);
    return pow((kPqC1 + kPqC2 * y) / (1.0 + kPqC3 * y), kPqM2);
}
/* BT.2100 HLG inverse OETF (color.cpp:81-84). */
static double hlg_inv_oetf(double e) {
    const double a = 0.17883277, b = 1.0 - 4.0 * a, c = 0.5 - a * log(4.0 * a);
    return e <= 0.5 ? e * e / 3.0 : (exp((e - c) / a) + b) / 12.0;
}

/* ColorConverter's constructor (color.cpp:296-372): the EOTF table and the parameter block
 * (color_params.h) for every tone map, tone norm and target gamut. Enumerations as in
 * ColorParams: tone_map 0 signal, 1 shoulder, 2 BT.2390, 3 none; tone_norm 0 channel,
 * 1 max(R,G,B), 2 luminance; gamut 0 P3-D65, 1 DCI-P3, 2 none. */
void or_color_setup(const or_color_cfg* c, float* lut, float* cp) {
    const int hdr = c->transfer == 2 || c->transfer == 3, passthrough = c->primaries == 4;
    for (int i = 0; i < 65536; ++i) {
        const double e = i / 65535.0;
        double nits = 0;
        switch (c->transfer) {
        case 0: nits = pow(e, 2.4) * c->sdr_white; break;
        case 1: nits = (e <= 0.04045 ? e / 12.92 : pow((e + 0.055) / 1.055, 2.4)) * c->sdr_white; break;
        case 2: nits = c->tone_map == 0 ? pow(e, 2.4) * c->sdr_white : pq_eotf_d(e); break;
        case 3: nits = c->tone_map == 0 ? pow(e, 2.4) * c->sdr_white : 1000.0 * pow(hlg_inv_oetf(e), 1.2); break;
        case 4: nits = e * c->sdr_white; break;
        case 5: nits = pow(e, 2.6) * 52.37; break;
        }
        lut[i] = (float)nits;
    }
    memset(cp, 0, 48 * sizeof(float));
    double M[9];
    rgb_to_xyz_matrix(c->primaries, M);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) cp[4 * r + k] = (float)M[3 * r + k];
    cp[17] = (float)c->cinema_peak;
    if (hdr && !passthrough && c->tone_map == 1) {
        const double gain = c->cinema_ref_white / c->cinema_peak;
        double knee = clampd(c->knee, 0.05, 0.99);
        const double vmax = gain * c->hdr_peak / c->hdr_ref_white;
        if (vmax <= 1.0) knee = 1.0;
        cp[12] = 1.f;
        cp[13] = (float)c->hdr_ref_white;
        cp[14] = (float)gain;
        cp[15] = (float)knee;
        cp[16] = (float)maxd(1e-3, (vmax - knee) / (1.0 - knee));
    } else if (hdr && !passthrough && c->tone_map == 2 && c->hdr_peak > c->cinema_peak) {
        cp[12] = 2.f;
        cp[39] = (float)pq_inv_d(c->hdr_peak);
        cp[40] = (float)(pq_inv_d(c->cinema_peak) / pq_inv_d(c->hdr_peak));
        const float ks = 1.5f * cp[40] - 0.5f;
        cp[41] = 0.f < ks ? ks : 0.f;
    }
    cp[42] = c->tone_norm == 0 ? 2.f : c->tone_norm == 1 ? 1.f : 0.f;
    if (!passthrough) {
        double T[9], Ti[9];
        rgb_to_xyz_matrix(c->gamut == 1 ? 3 : 2, T);
        invert3(T, Ti);
        for (int i = 0; i < 9; ++i) {
            cp[20 + i] = (float)(Ti[i] / c->cinema_peak);
            cp[29 + i] = (float)(T[i] * c->cinema_peak);
        }
        if (c->gamut != 2) {
            cp[18] = 1.f;
            cp[19] = (float)clampd(c->gamut_threshold, 0.0, 0.99);
        }
    }
    cp[38] = (float)c->lut3d_size;
}

/* cm_pq_eotf / cm_pq_inv in float (color_math.h:27-40). */
static float pq_eotf_f(float e) {
    const float m1 = 2610.0f / 16384.0f, m2 = 2523.0f / 4096.0f * 128.0f, c1 = 3424.0f / 4096.0f,
                c2 = 2413.0f / 4096.0f * 32.0f, c3 = 2392.0f / 4096.0f * 32.0f;
    const float p = powf(fminf(fmaxf(e, 0.0f), 1.0f), 1.0f / m2);
    return 10000.0f * powf(fmaxf(p - c1, 0.0f) / (c2 - c3 * p), 1.0f / m1);
}
static float pq_inv_f(float nits) {
    const float m1 = 2610.0f / 16384.0f, m2 = 2523.0f / 4096.0f * 128.0f, c1 = 3424.0f / 4096.0f,
                c2 = 2413.0f / 4096.0f * 32.0f, c3 = 2392.0f / 4096.0f * 32.0f;
    const float y = powf(fminf(fmaxf(nits / 10000.0f, 0.0f), 1.0f), m1);
    return powf((c1 + c2 * y) / (1.0f + c3 * y), m2);
}

/* cm_tone (color_math.h:42-65): target luminance for source luminance y. */
static float tone(const float* cp, float y) {
    if (y <= 0.0f) return 0.0f;
    if (cp[12] > 1.5f) {
        const float srcMax = cp[39], ks = cp[41];
        const float e1 = pq_inv_f(y) / srcMax;
        if (e1 < ks) return y;
        const float t = (fminf(e1, 1.0f) - ks) / (1.0f - ks);
        const float t2 = t * t, t3 = t2 * t;
        const float e2 = (2.0f * t3 - 3.0f * t2 + 1.0f) * ks + (t3 - 2.0f * t2 + t) * (1.0f - ks) +
                         (-2.0f * t3 + 3.0f * t2) * cp[40];
        return pq_eotf_f(e2 * srcMax);
    }
    float v = cp[14] * y / cp[13];
    const float k = cp[15];
    if (v > k) {
        const float t = (v - k) / (1.0f - k);
        const float T = cp[16];
        v = k + (1.0f - k) * fminf(t * (1.0f + t / (T * T)) / (1.0f + t), 1.0f);
    }
    return v * cp[17];
}

/* cm_gamut (color_math.h:67-105): toward neutral grey of the same luminance until every
 * target-gamut channel lies in [0, 1], compressed smoothly beyond the threshold. */
static void gamut(const float* cp, float* X, float* Y, float* Z) {
    float r = cp[20] * *X + cp[21] * *Y + cp[22] * *Z;
    float g = cp[23] * *X + cp[24] * *Y + cp[25] * *Z;
    float b = cp[26] * *X + cp[27] * *Y + cp[28] * *Z;
    const float n = *Y / cp[17];
    if (n >= 1.0f) {
        r = g = b = 1.0f;
    } else {
        float amax = 1.0f;
        const float d0 = r - n, d1 = g - n, d2 = b - n;
        if (d0 > 0.0f) amax = fminf(amax, (1.0f - n) / d0);
        if (d0 < 0.0f) amax = fminf(amax, -n / d0);
        if (d1 > 0.0f) amax = fminf(amax, (1.0f - n) / d1);
        if (d1 < 0.0f) amax = fminf(amax, -n / d1);
        if (d2 > 0.0f) amax = fminf(amax, (1.0f - n) / d2);
        if (d2 < 0.0f) amax = fminf(amax, -n / d2);
        const float dist = 1.0f / fmaxf(amax, 1e-6f);
        const float th = cp[19];
        float alpha = 1.0f;
        if (dist > th) {
            const float x = (dist - th) / (1.0f - th);
            const float y = x / (1.0f + x);
            alpha = (th + (1.0f - th) * y) / dist;
        }
        r = n + alpha * d0;
        g = n + alpha * d1;
        b = n + alpha * d2;
    }
    *X = cp[29] * r + cp[30] * g + cp[31] * b;
    *Y = cp[32] * r + cp[33] * g + cp[34] * b;
    *Z = cp[35] * r + cp[36] * g + cp[37] * b;
}

/* cm_rgb_to_xyz12 (color_math.h:114-179) for pixels given as 16-bit codes (code / 65535,
 * color.cpp:400-406): the 3D LUT route or the managed route with tone map and gamut. */
void or_color_pixels(const float* lut, const float* cp, const float* lut3, int n, const uint16_t* rgb, int* out) {
    for (int p = 0; p < n; ++p) {
        const float r = rgb[3 * p] / 65535.f, g = rgb[3 * p + 1] / 65535.f, b = rgb[3 * p + 2] / 65535.f;
        int* o = out + 3 * p;
        const int ln = (int)cp[38];
        if (ln > 1) {
            const float fr = fminf(fmaxf(r, 0.0f), 1.0f) * (float)(ln - 1);
            const float fg = fminf(fmaxf(g, 0.0f), 1.0f) * (float)(ln - 1);
            const float fb = fminf(fmaxf(b, 0.0f), 1.0f) * (float)(ln - 1);
            const int i0 = imin((int)fr, ln - 2), j0 = imin((int)fg, ln - 2), k0 = imin((int)fb, ln - 2);
            const float ar = fr - (float)i0, ag = fg - (float)j0, ab = fb - (float)k0;
#define AT(i, j, k, c) lut3[(((k) * ln + (j)) * ln + (i)) * 3 + (c)]
            for (int c = 0; c < 3; ++c) {
                const float c00 = AT(i0, j0, k0, c) * (1.0f - ar) + AT(i0 + 1, j0, k0, c) * ar;
                const float c10 = AT(i0, j0 + 1, k0, c) * (1.0f - ar) + AT(i0 + 1, j0 + 1, k0, c) * ar;
                const float c01 = AT(i0, j0, k0 + 1, c) * (1.0f - ar) + AT(i0 + 1, j0, k0 + 1, c) * ar;
                const float c11 = AT(i0, j0 + 1, k0 + 1, c) * (1.0f - ar) + AT(i0 + 1, j0 + 1, k0 + 1, c) * ar;
                const float v = (c00 * (1.0f - ag) + c10 * ag) * (1.0f - ab) + (c01 * (1.0f - ag) + c11 * ag) * ab;
                o[c] = (int)floorf(fminf(fmaxf(v, 0.0f), 1.0f) * 4095.0f + 0.5f);
            }
#undef AT
            continue;
        }
        const float R = lut[lut_index(r)], G = lut[lut_index(g)], B = lut[lut_index(b)];
        float X = cp[0] * R + cp[1] * G + cp[2] * B;
        float Y = cp[4] * R + cp[5] * G + cp[6] * B;
        float Z = cp[8] * R + cp[9] * G + cp[10] * B;
        if (cp[12] > 0.5f && Y > 0.0f) {
            float tr = cp[20] * X + cp[21] * Y + cp[22] * Z;
            float tg = cp[23] * X + cp[24] * Y + cp[25] * Z;
            float tb = cp[26] * X + cp[27] * Y + cp[28] * Z;
            const float pk = cp[17];
            if (cp[42] > 1.5f) {
                const float mx = fmaxf(tr, fmaxf(tg, tb)) * pk;
                const float sneg = mx > 0.0f ? tone(cp, mx) / mx : 1.0f;
                tr = tr > 0.0f ? tone(cp, tr * pk) / pk : tr * sneg;
                tg = tg > 0.0f ? tone(cp, tg * pk) / pk : tg * sneg;
                tb = tb > 0.0f ? tone(cp, tb * pk) / pk : tb * sneg;
                X = cp[29] * tr + cp[30] * tg + cp[31] * tb;
                Y = cp[32] * tr + cp[33] * tg + cp[34] * tb;
                Z = cp[35] * tr + cp[36] * tg + cp[37] * tb;
            } else {
                const float m = cp[42] > 0.5f ? fmaxf(Y, fmaxf(tr, fmaxf(tg, tb)) * pk) : Y;
                const float s = tone(cp, m) / m;
                X *= s;
                Y *= s;
                Z *= s;
            }
        }
        if (cp[18] > 0.5f) gamut(cp, &X, &Y, &Z);
        o[0] = (int)encode_code(X);
        o[1] = (int)encode_code(Y);
        o[2] = (int)encode_code(Z);
    }
}

uint16_t or_xyz12_to_16(uint32_t v) {
    if (v > 4095u) v = 4095u;
    return (uint16_t)((v * 65535u + 2047u) / 4095u);
}

/* lround for v in [0, 65535] (color.cpp:136-139). */
static uint16_t round16(float v) {
    const int r = (int)v;
    return (uint16_t)(r + ((v - (float)r) >= 0.5f ? 1 : 0));
}

static inline float sample(const void* src, int bits, size_t i) {
    return bits > 8 ? (float)((const uint16_t*)src)[i] : (float)((const uint8_t*)src)[i];
}

static inline float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }

/* planarToRgb48Row over the whole container (color.cpp:141-203): bilinear chroma at the
 * chroma siting, range expansion, Kr/Kb matrix, rounding to 16 bits; black outside the
 * active picture. */
void or_planar_to_rgb48(const void* src, int bits, int aw, int ah, int sx, int sy, int rgb, int full_range,
                        int chroma, double kr, double kb, int width, int height, int ox, int oy, uint16_t* out) {
    const int cw = (aw + sx) >> sx, ch = (ah + sy) >> sy;
    const size_t u_off = (size_t)aw * ah, v_off = u_off + (size_t)cw * ch;
    const float maxv = (float)((1 << bits) - 1), scale = (float)(1 << (bits - 8));
    const double kg = 1.0 - kr - kb;
    const float crR = (float)(2 * (1 - kr)), cbB = (float)(2 * (1 - kb));
    const float cbG = (float)(2 * kb * (1 - kb) / kg), crG = (float)(2 * kr * (1 - kr) / kg);
    memset(out, 0, sizeof(uint16_t) * 3 * (size_t)width * height);
    for (int ay = 0; ay < ah; ++ay) {
        float cyp = !sy ? (float)ay : chroma == 2 ? ay * 0.5f : (ay - 0.5f) * 0.5f;
        cyp = cyp < 0.f ? 0.f : (cyp > (float)(ch - 1) ? (float)(ch - 1) : cyp);
        const int y0 = (int)cyp, y1 = imin(y0 + 1, ch - 1);
        const float fy = cyp - y0;
        uint16_t* o = out + ((size_t)(ay + oy) * width + ox) * 3;
        for (int x = 0; x < aw; ++x, o += 3) {
            float cxp = !sx ? (float)x : chroma == 1 ? (x - 0.5f) * 0.5f : x * 0.5f;
            cxp = cxp < 0.f ? 0.f : (cxp > (float)(cw - 1) ? (float)(cw - 1) : cxp);
            const int x0 = (int)cxp, x1 = imin(x0 + 1, cw - 1);
            const float fx = cxp - x0;
            float uv[2];
            for (int k = 0; k < 2; ++k) {
                const size_t off = k ? v_off : u_off;
                const float a = sample(src, bits, off + (size_t)y0 * cw + x0);
                const float b = sample(src, bits, off + (size_t)y0 * cw + x1);
                const float c = sample(src, bits, off + (size_t)y1 * cw + x0);
                const float d = sample(src, bits, off + (size_t)y1 * cw + x1);
                uv[k] = (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
            }
            float Y = sample(src, bits, (size_t)ay * aw + x), U = uv[0], V = uv[1];
            float f[3];
            if (rgb) { /* planes G', B', R' */
                f[0] = V / maxv;
                f[1] = Y / maxv;
                f[2] = U / maxv;
            } else {
                if (full_range) {
                    Y /= maxv;
                    U = (U - 128.f * scale) / maxv;
                    V = (V - 128.f * scale) / maxv;
                } else {
                    Y = (Y - 16.f * scale) / (219.f * scale);
                    U = (U - 128.f * scale) / (224.f * scale);
                    V = (V - 128.f * scale) / (224.f * scale);
                }
                f[0] = Y + crR * V;
                f[1] = Y - cbG * U - crG * V;
                f[2] = Y + cbB * U;
            }
            for (int c = 0; c < 3; ++c) o[c] = round16(clamp01(f[c]) * 65535.f);
        }
    }
}

static inline void ict(float x, float y, float z, float* c0, float* c1, float* c2) {
    *c0 = (float)ICT_FWD[0][0] * x + (float)ICT_FWD[0][1] * y + (float)ICT_FWD[0][2] * z;
    *c1 = (float)ICT_FWD[1][0] * x + (float)ICT_FWD[1][1] * y + (float)ICT_FWD[1][2] * z;
    *c2 = (float)ICT_FWD[2][0] * x + (float)ICT_FWD[2][1] * y + (float)ICT_FWD[2][2] * z;
}

/* ColorConverter::convertFrame (color.cpp:413-430): codes, DC shift, forward ICT. */
void or_convert_frame(const float* lut, const float* cp, const uint16_t* rgb, int width, int height,
                      float* planes, uint16_t* xyz12) {
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; ++i) {
        int c[3];
        or_color_pixel(lut, cp, rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], c);
        if (xyz12)
            for (int k = 0; k < 3; ++k) xyz12[3 * i + k] = (uint16_t)c[k];
        ict((float)(c[0] - 2048), (float)(c[1] - 2048), (float)(c[2] - 2048), &planes[i], &planes[n + i],
            &planes[2 * n + i]);
    }
}

void or_xyz_planar_to_ict(const uint16_t* xyz, int width, int height, float* planes) {
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; ++i)
        ict((float)((int)xyz[i] - 2048), (float)((int)xyz[n + i] - 2048), (float)((int)xyz[2 * n + i] - 2048),
            &planes[i], &planes[n + i], &planes[2 * n + i]);
}

/* ======================================================================== quantiser */
/* Dead-zone quantiser with 6 fraction bits (quant.h:20-25). */
uint32_t or_quantize_coeff(float v, float inv_delta, uint32_t max_mag) {
    const float a = v < 0 ? -v : v;
    const float q = a * inv_delta * 64.0f;
    const uint32_t m = q >= (float)max_mag ? max_mag : (uint32_t)q;
    return m | ((v < 0 && m >= 64u) ? 0x80000000u : 0u);
}

void or_quantize_block(const float* plane, int stride, int x0, int y0, int w, int h, float delta, int mb,
                       uint32_t* out) {
    const float inv = 1.0f / delta;
    const uint32_t max_mag = ((1u << mb) << 6) - 1;
    memset(out, 0, 1024 * sizeof(uint32_t));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) out[y * 32 + x] = or_quantize_coeff(plane[(size_t)(y0 + y) * stride + x0 + x], inv, max_mag);
}

/* ======================================================================== MQ coder */
/* ISO/IEC 15444-1 Table C.2: Qe, next index after MPS / LPS, MPS switch. */
static const uint16_t MQ_QE[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801, 0x3001, 0x2401,
    0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401, 0x3001, 0x2801, 0x2401, 0x2201,
    0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101, 0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1,
    0x0221, 0x0141, 0x0111, 0x0085, 0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
static const uint8_t MQ_NMPS[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                    17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                    33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
static const uint8_t MQ_NLPS[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                    15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                    30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
static const uint8_t MQ_SW[47] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                  0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

enum { CX_SC = 9, CX_MR = 14, CX_RL = 17, CX_UNI = 18, NCX = 19 };

typedef struct {
    uint32_t A, C;
    int CT, bp;   /* bp indexes buf; buf[-1] is a scratch byte before the codeword */
    uint8_t* buf; /* caller guarantees buf[-1] is addressable */
    uint32_t cap;
    int overflow;
    uint8_t I[NCX], mps[NCX];
} mq_enc;

typedef struct {
    int bp;
    uint32_t B, C, A;
    int CT;
} mq_snap;

/* INITENC plus the context states of Table D.7. */
static void mq_start(mq_enc* e, uint8_t* buf, uint32_t cap) {
    e->A = 0x8000u;
    e->C = 0;
    e->CT = 12;
    e->bp = -1;
    e->buf = buf;
    e->cap = cap;
    e->overflow = 0;
    buf[-1] = 0;
    memset(e->I, 0, sizeof e->I);
    memset(e->mps, 0, sizeof e->mps);
    e->I[0] = 4;
    e->I[CX_RL] = 3;
    e->I[CX_UNI] = 46;
}

static inline void mq_put(mq_enc* e, int i, uint32_t v) {
    if (i < (int)e->cap) e->buf[i] = (uint8_t)v;
    else e->overflow = 1;
}
static inline uint32_t mq_cur(const mq_enc* e) { return e->bp < (int)e->cap ? e->buf[e->bp] : 0; }

/* BYTEOUT (C.2.7) with bit stuffing after 0xFF and carry propagation into the byte at bp. */
static void mq_byteout(mq_enc* e) {
    uint32_t B = mq_cur(e);
    if (B != 0xFFu && (e->C & 0x8000000u)) { /* carry */
        ++B;
        mq_put(e, e->bp, B);
        if (B == 0xFFu) e->C &= 0x7FFFFFFu;
    }
    ++e->bp;
    if (B == 0xFFu) {
        mq_put(e, e->bp, e->C >> 20);
        e->C &= 0xFFFFFu;
        e->CT = 7;
    } else {
        mq_put(e, e->bp, (e->C >> 19) & 0xFFu);
        e->C &= 0x7FFFFu;
        e->CT = 8;
    }
}

static void mq_renorm(mq_enc* e) {
    do {
        e->A <<= 1;
        e->C <<= 1;
        if (--e->CT == 0) mq_byteout(e);
    } while (!(e->A & 0x8000u));
}

/* ENCODE = CODEMPS / CODELPS (C.2.5, C.2.6). */
static void mq_code(mq_enc* e, int cx, uint32_t d) {
    const int i = e->I[cx];
    const uint32_t qe = MQ_QE[i];
    e->A -= qe;
    if (d == e->mps[cx]) {
        if (e->A & 0x8000u) {
            e->C += qe;
            return;
        }
        if (e->A < qe) e->A = qe;
        else e->C += qe;
        e->I[cx] = MQ_NMPS[i];
    } else {
        if (e->A < qe) e->C += qe;
        else e->A = qe;
        if (MQ_SW[i]) e->mps[cx] ^= 1;
        e->I[cx] = MQ_NLPS[i];
    }
    mq_renorm(e);
}

static mq_snap mq_snapshot(const mq_enc* e) {
    mq_snap s = {e->bp, e->bp >= 0 ? mq_cur(e) : 0u, e->C, e->A, e->CT};
    return s;
}

/* FLUSH (C.2.9): SETBITS, two BYTEOUTs; a final 0xFF is not part of the codeword. */
static uint32_t mq_flush(mq_enc* e) {
    const uint32_t tempc = e->C + e->A;
    e->C |= 0xFFFFu;
    if (e->C >= tempc) e->C -= 0x8000u;
    e->C <<= e->CT;
    mq_byteout(e);
    e->C <<= e->CT;
    mq_byteout(e);
    return mq_cur(e) != 0xFFu ? (uint32_t)(e->bp + 1) : (uint32_t)e->bp;
}

/* Truncation length after a pass (mq_encoder.h:146-169): the shortest prefix L >= min_len
 * of the final codeword (never ending in 0xFF) such that the prefix followed by the
 * decoder's all-ones fill lies inside the coding interval [low, low + A) the encoder had
 * at the end of the pass. Byte i of the codeword weighs 2^(43 - CT - rel), where rel
 * counts the bits of the bytes from s.bp up to i (7 after a 0xFF, else 8); units are
 * 2^-16 of the snapshot's C register LSB. */
static uint32_t mq_trunc_len(const mq_snap* s, const uint8_t* buf, uint32_t len, uint32_t min_len) {
    const int base = 43 - s->CT;
    const uint64_t low = ((uint64_t)s->B << base) + ((uint64_t)s->C << 16);
    const uint64_t high = low + ((uint64_t)s->A << 16);
    uint64_t prefix = 0;
    int rel = 0;
    for (int i = s->bp; i < (int)len; ++i) {
        const uint32_t byte = i >= 0 ? buf[i] : 0u;
        const int sh = base - rel;
        if (sh < 0) break;
        prefix += (uint64_t)byte << sh;
        const uint64_t tail = prefix + ((uint64_t)1 << sh);
        if ((uint32_t)(i + 1) >= min_len && byte != 0xFFu && tail <= high && tail > low) return (uint32_t)(i + 1);
        rel += byte == 0xFFu ? 7 : 8;
    }
    return len;
}

/* ======================================================================== Tier-1 */
/* Significance context, Table D.1. h, v: significant horizontal / vertical neighbours,
 * d: diagonal. orient 1 (HL) swaps the roles of h and v; 3 (HH) keys on d. */
static int zc_context(int h, int v, int d, int orient) {
    if (orient == 1) {
        const int t = h;
        h = v;
        v = t;
    }
    if (orient == 3) {
        const int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv ? 4 : 3);
        return hv >= 2 ? 2 : hv;
    }
    if (h == 2) return 8;
    if (h == 1) return v ? 7 : (d ? 6 : 5);
    if (v) return v == 2 ? 4 : 3;
    return d >= 2 ? 2 : d;
}

/* Sign context and XOR bit, Table D.3, from the clamped horizontal and vertical
 * contributions. */
static void sc_context(int hc, int vc, int* cx, int* xr) {
    static const int ctx[3][3] = {{13, 12, 11}, {10, 9, 10}, {11, 12, 13}}; /* [hc+1][vc+1] */
    static const int xb[3][3] = {{1, 1, 1}, {1, 0, 0}, {0, 0, 0}};
    *cx = ctx[hc + 1][vc + 1];
    *xr = xb[hc + 1][vc + 1];
}

enum { T_SIG = 1, T_SGN = 2, T_VIS = 4, T_REF = 8 };

typedef struct {
    const uint32_t* c;
    int w, h, orient, stride;
    uint8_t* f; /* (w + 2) x (h + 2) flags with a zero border */
    mq_enc mq;
    int32_t dist;
    int track;
    uint32_t hist[NCX][2];
    uint8_t* syms;
    uint32_t sym_cap, nsyms;
} t1c;

static inline uint8_t* F(t1c* t, int x, int y) { return &t->f[(size_t)(y + 1) * t->stride + x + 1]; }
static inline uint32_t MAG(const t1c* t, int x, int y) { return t->c[y * 32 + x] & 0x7FFFFFFFu; }
static inline int sigf(t1c* t, int x, int y) { return *F(t, x, y) & T_SIG; }
static inline int chi(t1c* t, int x, int y) {
    const uint8_t f = *F(t, x, y);
    return (f & T_SIG) ? ((f & T_SGN) ? -1 : 1) : 0;
}

static void neighbours(t1c* t, int x, int y, int* h, int* v, int* d) {
    *h = sigf(t, x - 1, y) + sigf(t, x + 1, y);
    *v = sigf(t, x, y - 1) + sigf(t, x, y + 1);
    *d = sigf(t, x - 1, y - 1) + sigf(t, x + 1, y - 1) + sigf(t, x - 1, y + 1) + sigf(t, x + 1, y + 1);
}

static void emit(t1c* t, int cx, uint32_t d) {
    mq_code(&t->mq, cx, d);
    t->hist[cx][d]++;
    if (t->syms) {
        if (t->nsyms < t->sym_cap) t->syms[t->nsyms] = (uint8_t)((cx << 1) | d);
        ++t->nsyms;
    }
}

static uint32_t bit_at(const t1c* t, int x, int y, int p) { return (MAG(t, x, y) >> (p + 6)) & 1u; }

/* Codes the sign of (x, y), makes it significant and credits its distortion reduction
 * (t1_ref.cpp:166-190; t1_common.h:51). */
static void become_significant(t1c* t, int x, int y, int p) {
    int hc = chi(t, x - 1, y) + chi(t, x + 1, y), vc = chi(t, x, y - 1) + chi(t, x, y + 1);
    hc = hc < -1 ? -1 : (hc > 1 ? 1 : hc);
    vc = vc < -1 ? -1 : (vc > 1 ? 1 : vc);
    int cx, xr;
    sc_context(hc, vc, &cx, &xr);
    const uint32_t s = (*F(t, x, y) & T_SGN) ? 1u : 0u;
    emit(t, cx, s ^ (uint32_t)xr);
    *F(t, x, y) |= T_SIG;
    const int32_t w = (int32_t)((MAG(t, x, y) >> p) & 0x7Fu);
    t->dist += 192 * w - 9216;
}

/* Significance propagation pass (D.3.1; t1_ref.cpp:200-215). */
static void pass_spp(t1c* t, int p) {
    for (int y0 = 0; y0 < t->h; y0 += 4)
        for (int x = 0; x < t->w; ++x)
            for (int y = y0; y < imin(y0 + 4, t->h); ++y) {
                if (sigf(t, x, y)) continue;
                int h, v, d;
                neighbours(t, x, y, &h, &v, &d);
                if (!(h | v | d)) continue;
                const uint32_t b = bit_at(t, x, y, p);
                emit(t, zc_context(h, v, d, t->orient), b);
                *F(t, x, y) |= T_VIS;
                if (b) become_significant(t, x, y, p);
            }
}

/* Magnitude refinement pass (D.3.3; t1_ref.cpp:217-232). */
static void pass_mrp(t1c* t, int p) {
    for (int y0 = 0; y0 < t->h; y0 += 4)
        for (int x = 0; x < t->w; ++x)
            for (int y = y0; y < imin(y0 + 4, t->h); ++y) {
                uint8_t* f = F(t, x, y);
                if (!(*f & T_SIG) || (*f & T_VIS)) continue;
                int h, v, d;
                neighbours(t, x, y, &h, &v, &d);
                const int cx = (*f & T_REF) ? 16 : ((h | v | d) ? 15 : 14);
                emit(t, cx, bit_at(t, x, y, p));
                *f |= T_REF;
                const int32_t w = (int32_t)((MAG(t, x, y) >> p) & 0x7Fu);
                t->dist += w >= 64 ? 64 * w - 5120 : 3072 - 64 * w;
            }
}

/* Cleanup pass with run-length mode (D.3.4; t1_ref.cpp:234-270). */
static void pass_cup(t1c* t, int p) {
    for (int y0 = 0; y0 < t->h; y0 += 4) {
        const int rows = imin(4, t->h - y0);
        for (int x = 0; x < t->w; ++x) {
            int r = 0;
            if (rows == 4) {
                int quiet = 1;
                for (int k = 0; k < 4 && quiet; ++k) {
                    int h, v, d;
                    neighbours(t, x, y0 + k, &h, &v, &d);
                    if ((*F(t, x, y0 + k) & (T_SIG | T_VIS)) || (h | v | d)) quiet = 0;
                }
                if (quiet) {
                    int first = -1;
                    for (int k = 0; k < 4; ++k)
                        if (bit_at(t, x, y0 + k, p)) {
                            first = k;
                            break;
                        }
                    emit(t, CX_RL, first >= 0);
                    if (first < 0) continue;
                    emit(t, CX_UNI, (uint32_t)first >> 1);
                    emit(t, CX_UNI, (uint32_t)first & 1u);
                    become_significant(t, x, y0 + first, p);
                    r = first + 1;
                }
            }
            for (; r < rows; ++r) {
                const int y = y0 + r;
                if (*F(t, x, y) & (T_SIG | T_VIS)) continue;
                int h, v, d;
                neighbours(t, x, y, &h, &v, &d);
                const uint32_t b = bit_at(t, x, y, p);
                emit(t, zc_context(h, v, d, t->orient), b);
                if (b) become_significant(t, x, y, p);
            }
        }
    }
}

/* Entropy estimate (bytes) of the plane's symbols from per-context counts
 * (t1_ref.cpp:128-140); clears the counts. */
static float plane_bytes(t1c* t) {
    float bits = 0.f;
    for (int cx = 0; cx < NCX; ++cx) {
        const float n0 = (float)t->hist[cx][0], n1 = (float)t->hist[cx][1], n = n0 + n1;
        if (n > 0.f) {
            const float a = n0 * log2f((n0 > 0.5f ? n0 : 0.5f) / n);
            const float b = n1 * log2f((n1 > 0.5f ? n1 : 0.5f) / n);
            bits += -(a + b) + 2.f;
        }
        t->hist[cx][0] = t->hist[cx][1] = 0;
    }
    return bits * 0.125f + 1.f;
}

/* Squared error (units of delta^2) left once bit-planes >= p are known, mid-point
 * reconstruction, summed row by row (t1_ref.cpp:142-154). */
static float residual(const t1c* t, int p) {
    float e = 0.f;
    const float step = exp2f((float)p);
    for (int y = 0; y < t->h; ++y)
        for (int x = 0; x < t->w; ++x) {
            const uint32_t m = MAG(t, x, y);
            const float v = (float)m * (1.f / 64.f);
            const uint32_t q = m >> (p + 6);
            const float rec = q == 0 ? 0.f : ((float)q + 0.5f) * step;
            e += (v - rec) * (v - rec);
        }
    return e;
}

int or_t1_encode(const uint32_t* coeffs, int w, int h, int orient, int max_passes, float slope_thresh,
                 float dist_scale, uint8_t* out, uint32_t out_cap, or_t1_result* res, uint8_t* syms,
                 uint32_t sym_cap, uint32_t* nsyms, uint32_t* pass_end) {
    memset(res, 0, sizeof *res);
    uint32_t maxmag = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const uint32_t m = (coeffs[y * 32 + x] & 0x7FFFFFFFu) >> 6;
            if (m > maxmag) maxmag = m;
        }
    int nbps = 0;
    while (nbps < 32 && (maxmag >> nbps)) ++nbps;
    res->nbps = nbps;
    if (nsyms) *nsyms = 0;
    if (!nbps) return 0;
    if (3 * nbps - 2 > OR_MAX_PASSES) return -1;
    t1c t;
    memset(&t, 0, sizeof t);
    t.c = coeffs;
    t.w = w;
    t.h = h;
    t.orient = orient;
    t.stride = w + 2;
    t.f = (uint8_t*)calloc((size_t)(w + 2) * (h + 2), 1);
    t.track = slope_thresh > 0.f;
    t.syms = syms;
    t.sym_cap = sym_cap;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (coeffs[y * 32 + x] & 0x80000000u) *F(&t, x, y) = T_SGN;
    uint8_t* buf = (uint8_t*)malloc((size_t)out_cap + 1);
    mq_start(&t.mq, buf + 1, out_cap);
    mq_snap snaps[OR_MAX_PASSES];
    int np = 0;
    for (int p = nbps - 1; p >= 0 && np < max_passes; --p) {
        int32_t plane_dist = 0;
        for (int type = (p == nbps - 1 ? 2 : 0); type < 3 && np < max_passes; ++type) {
            t.dist = 0;
            if (type == 0) pass_spp(&t, p);
            else if (type == 1) pass_mrp(&t, p);
            else pass_cup(&t, p);
            snaps[np] = mq_snapshot(&t.mq);
            res->pass_dist[np] = t.dist;
            plane_dist += t.dist;
            if (pass_end) pass_end[np] = t.nsyms;
            ++np;
        }
        for (size_t i = 0; i < (size_t)(w + 2) * (h + 2); ++i) t.f[i] &= (uint8_t)~T_VIS;
        if (t.track) {
            /* early termination (t1_ref.cpp:156-162) */
            const float bytes = plane_bytes(&t);
            if (p > 0) {
                const float dabs = (float)plane_dist * exp2f((float)(2 * p)) * (1.f / 4096.f) * dist_scale;
                if (dabs < slope_thresh * bytes && residual(&t, p) * dist_scale < slope_thresh * bytes) break;
            }
        }
    }
    const uint32_t len = mq_flush(&t.mq);
    int rc = t.mq.overflow ? -1 : 0;
    if (!rc) {
        res->num_passes = np;
        res->len = len;
        uint32_t prev = 1;
        for (int k = 0; k < np; ++k) {
            const uint32_t L = k == np - 1 ? len : mq_trunc_len(&snaps[k], buf + 1, len, prev);
            res->pass_len[k] = L;
            if (L > prev) prev = L;
        }
        memcpy(out, buf + 1, len);
    }
    if (nsyms) *nsyms = t.nsyms;
    free(buf);
    free(t.f);
    return rc;
}

/* ======================================================================== decoder */
/* Independent Tier-1 decoder (Annexes C.3 and D) used to check the encoder. */
typedef struct {
    const uint8_t* d;
    uint32_t len, bp;
    uint32_t A, C;
    int CT;
    uint8_t I[NCX], mps[NCX];
} mq_dec;

static inline uint32_t dbyte(const mq_dec* m, uint32_t i) { return i < m->len ? m->d[i] : 0xFFu; }

static void mq_bytein(mq_dec* m) { /* C.3.4 */
    if (dbyte(m, m->bp) == 0xFFu) {
        if (dbyte(m, m->bp + 1) > 0x8Fu) {
            m->C += 0xFF00u;
            m->CT = 8;
        } else {
            ++m->bp;
            m->C += dbyte(m, m->bp) << 9;
            m->CT = 7;
        }
    } else {
        ++m->bp;
        m->C += dbyte(m, m->bp) << 8;
        m->CT = 8;
    }
}

static void mq_dec_start(mq_dec* m, const uint8_t* d, uint32_t len) { /* INITDEC, C.3.5 */
    m->d = d;
    m->len = len;
    m->bp = 0;
    m->C = dbyte(m, 0) << 16;
    mq_bytein(m);
    m->C <<= 7;
    m->CT -= 7;
    m->A = 0x8000u;
    memset(m->I, 0, sizeof m->I);
    memset(m->mps, 0, sizeof m->mps);
    m->I[0] = 4;
    m->I[CX_RL] = 3;
    m->I[CX_UNI] = 46;
}

static void mq_renormd(mq_dec* m) {
    do {
        if (m->CT == 0) mq_bytein(m);
        m->A <<= 1;
        m->C <<= 1;
        --m->CT;
    } while (!(m->A & 0x8000u));
}

static uint32_t mq_decode(mq_dec* m, int cx) { /* DECODE, C.3.2 */
    const int i = m->I[cx];
    const uint32_t qe = MQ_QE[i];
    uint32_t d;
    m->A -= qe;
    if ((m->C >> 16) < qe) { /* LPS exchange */
        if (m->A < qe) {
            d = m->mps[cx];
            m->I[cx] = MQ_NMPS[i];
        } else {
            d = 1u - m->mps[cx];
            if (MQ_SW[i]) m->mps[cx] ^= 1;
            m->I[cx] = MQ_NLPS[i];
        }
        m->A = qe;
        mq_renormd(m);
    } else {
        m->C -= qe << 16;
        if (!(m->A & 0x8000u)) { /* MPS exchange */
            if (m->A < qe) {
                d = 1u - m->mps[cx];
                if (MQ_SW[i]) m->mps[cx] ^= 1;
                m->I[cx] = MQ_NLPS[i];
            } else {
                d = m->mps[cx];
                m->I[cx] = MQ_NMPS[i];
            }
            mq_renormd(m);
        } else {
            d = m->mps[cx];
        }
    }
    return d;
}

typedef struct {
    int w, h, orient, stride;
    uint8_t* f; /* T_SIG, T_SGN (decoded), T_VIS, T_REF */
    uint32_t* mag;
    uint32_t* word; /* optional: the reconstruction words of t1_dec.cpp, w x h */
    mq_dec m;
} t1d;

static inline uint8_t* DF(t1d* t, int x, int y) { return &t->f[(size_t)(y + 1) * t->stride + x + 1]; }
static inline int dsig(t1d* t, int x, int y) { return *DF(t, x, y) & T_SIG; }
static inline int dchi(t1d* t, int x, int y) {
    const uint8_t f = *DF(t, x, y);
    return (f & T_SIG) ? ((f & T_SGN) ? -1 : 1) : 0;
}
static void dnb(t1d* t, int x, int y, int* h, int* v, int* d) {
    *h = dsig(t, x - 1, y) + dsig(t, x + 1, y);
    *v = dsig(t, x, y - 1) + dsig(t, x, y + 1);
    *d = dsig(t, x - 1, y - 1) + dsig(t, x + 1, y - 1) + dsig(t, x - 1, y + 1) + dsig(t, x + 1, y + 1);
}
static void dsign(t1d* t, int x, int y, int p) {
    int hc = dchi(t, x - 1, y) + dchi(t, x + 1, y), vc = dchi(t, x, y - 1) + dchi(t, x, y + 1);
    hc = hc < -1 ? -1 : (hc > 1 ? 1 : hc);
    vc = vc < -1 ? -1 : (vc > 1 ? 1 : vc);
    int cx, xr;
    sc_context(hc, vc, &cx, &xr);
    const uint32_t s = mq_decode(&t->m, cx) ^ (uint32_t)xr;
    *DF(t, x, y) |= (uint8_t)(T_SIG | (s ? T_SGN : 0));
    t->mag[y * t->w + x] |= 1u << p;
    /* twice the magnitude so far plus the mid-point of the planes below: 2 * 2^p + 2^p */
    if (t->word) t->word[y * t->w + x] = (s << 31) | (3u << p);
}

static int t1_decode_core(const uint8_t* data, uint32_t len, int w, int h, int orient, int nbps, int passes,
                          uint32_t* mag, uint32_t* word, uint8_t* sign, uint8_t* sig) {
    t1d t;
    t.w = w;
    t.h = h;
    t.orient = orient;
    t.stride = w + 2;
    t.f = (uint8_t*)calloc((size_t)(w + 2) * (h + 2), 1);
    t.mag = mag;
    t.word = word;
    memset(mag, 0, sizeof(uint32_t) * (size_t)w * h);
    if (word) memset(word, 0, sizeof(uint32_t) * (size_t)w * h);
    mq_dec_start(&t.m, data, len);
    int k = 0;
    for (int p = nbps - 1; p >= 0 && k < passes; --p) {
        for (int type = (p == nbps - 1 ? 2 : 0); type < 3 && k < passes; ++type, ++k) {
            for (int y0 = 0; y0 < h; y0 += 4) {
                const int rows = imin(4, h - y0);
                for (int x = 0; x < w; ++x) {
                    int r = 0;
                    if (type == 2 && rows == 4) {
                        int quiet = 1;
                        for (int q = 0; q < 4 && quiet; ++q) {
                            int a, b, c;
                            dnb(&t, x, y0 + q, &a, &b, &c);
                            if ((*DF(&t, x, y0 + q) & (T_SIG | T_VIS)) || (a | b | c)) quiet = 0;
                        }
                        if (quiet) {
                            if (!mq_decode(&t.m, CX_RL)) continue;
                            int first = (int)mq_decode(&t.m, CX_UNI) << 1;
                            first |= (int)mq_decode(&t.m, CX_UNI);
                            dsign(&t, x, y0 + first, p);
                            r = first + 1;
                        }
                    }
                    for (; r < rows; ++r) {
                        const int y = y0 + r;
                        uint8_t* f = DF(&t, x, y);
                        int a, b, c;
                        dnb(&t, x, y, &a, &b, &c);
                        if (type == 0) {
                            if ((*f & T_SIG) || !(a | b | c)) continue;
                            *f |= T_VIS;
                            if (mq_decode(&t.m, zc_context(a, b, c, orient))) dsign(&t, x, y, p);
                        } else if (type == 1) {
                            if (!(*f & T_SIG) || (*f & T_VIS)) continue;
                            const int cx = (*f & T_REF) ? 16 : ((a | b | c) ? 15 : 14);
                            const uint32_t bit = mq_decode(&t.m, cx);
                            t.mag[y * w + x] |= bit << p;
                            /* the bit replaces the mid-point guess: +-2^p in the doubled value */
                            if (t.word) t.word[y * w + x] += bit ? (1u << p) : (0u - (1u << p));
                            *f |= T_REF;
                        } else {
                            if (*f & (T_SIG | T_VIS)) continue;
                            if (mq_decode(&t.m, zc_context(a, b, c, orient))) dsign(&t, x, y, p);
                        }
                    }
                }
            }
        }
        for (size_t i = 0; i < (size_t)(w + 2) * (h + 2); ++i) t.f[i] &= (uint8_t)~T_VIS;
    }
    if (sign && sig)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const uint8_t f = *DF(&t, x, y);
                sign[y * w + x] = (f & T_SGN) ? 1 : 0;
                sig[y * w + x] = (f & T_SIG) ? 1 : 0;
            }
    free(t.f);
    return 0;
}

int or_t1_decode(const uint8_t* data, uint32_t len, int w, int h, int orient, int nbps, int passes,
                 uint32_t* mag, uint8_t* sign, uint8_t* sig) {
    return t1_decode_core(data, len, w, h, orient, nbps, passes, mag, NULL, sign, sig);
}

/* t1DecodeBlock (t1_dec.cpp:175-178): the words of t1_dec.cpp:85-97, 117-127 for w x h
 * samples (row stride w). Passes past the last plane are not decoded. */
int or_t1_decode_words(const uint8_t* data, uint32_t len, int w, int h, int orient, int nbps, int passes,
                       uint32_t* words) {
    if (w < 1 || h < 1 || w > 32 || h > 32) return -1;
    uint32_t mag[1024];
    return t1_decode_core(data, len, w, h, orient, nbps, passes, mag, words, NULL, NULL);
}

/* ======================================================================== frame */
typedef struct {
    atomic_int next;
    int n;
    void (*fn)(void* ctx, int i);
    void* ctx;
} par_job;

static void* par_worker(void* arg) {
    par_job* j = (par_job*)arg;
    for (;;) {
        const int i = atomic_fetch_add(&j->next, 1);
        if (i >= j->n) break;
        j->fn(j->ctx, i);
    }
    return NULL;
}

static void par_for(int n, int threads, void (*fn)(void*, int), void* ctx) {
    par_job j;
    atomic_init(&j.next, 0);
    j.n = n;
    j.fn = fn;
    j.ctx = ctx;
    if (threads < 1) threads = 1;
    pthread_t tid[256];
    const int nt = imin(threads, 256) - 1;
    for (int i = 0; i < nt; ++i) pthread_create(&tid[i], NULL, par_worker, &j);
    par_worker(&j);
    for (int i = 0; i < nt; ++i) pthread_join(tid[i], NULL);
}

typedef struct {
    float* planes;
    int W, H, w, h;
    const or_band* bands;
    const or_cblk* cblks;
    uint32_t* coeffs;
    or_t1_result* res;
    uint8_t* bufs;
    float thr;
    int failed;
} frame_ctx;

static void job_cols(void* p, int i) {
    frame_ctx* c = (frame_ctx*)p;
    const int comp = i / c->w, x = i % c->w;
    float tmp[2160];
    fwd1d_f(c->planes + (size_t)comp * c->W * c->H + x, c->h, c->W, tmp);
}
static void job_rows(void* p, int i) {
    frame_ctx* c = (frame_ctx*)p;
    const int comp = i / c->h, y = i % c->h;
    float tmp[4096];
    fwd1d_f(c->planes + (size_t)comp * c->W * c->H + (size_t)y * c->W, c->w, 1, tmp);
}
#define BLOCK_BUF (1024 * 4 * 24 + 64)
static void job_block(void* p, int i) {
    frame_ctx* c = (frame_ctx*)p;
    const or_cblk* cb = &c->cblks[i];
    const or_band* b = &c->bands[cb->band];
    uint32_t* q = c->coeffs + (size_t)i * 1024;
    or_quantize_block(c->planes + (size_t)b->comp * c->W * c->H, c->W, b->px0 + cb->x0, b->py0 + cb->y0, cb->w, cb->h,
                      b->delta, b->mb, q);
    const float scale = (float)((double)b->delta * b->delta * b->weight);
    if (or_t1_encode(q, cb->w, cb->h, b->orient, OR_MAX_PASSES, c->thr, scale, c->bufs + (size_t)i * BLOCK_BUF, BLOCK_BUF,
                     &c->res[i], NULL, 0, NULL, NULL))
        c->failed = 1;
}

int64_t or_encode_frame_xyz(const uint16_t* xyz, int W, int H, int fourk, double base_step, float slope_thresh,
                            int threads, or_cblk_info* info, uint32_t* pass_len, int32_t* pass_dist, uint8_t* data,
                            uint64_t data_cap, uint32_t* coeffs, float* planes_out) {
    or_band bands[OR_MAX_BANDS];
    int nb = 0;
    const int n = or_layout_make(W, H, fourk, base_step, bands, &nb, NULL, 0);
    if (n < 0) return -1;
    or_cblk* cblks = (or_cblk*)malloc(sizeof(or_cblk) * (size_t)n);
    or_layout_make(W, H, fourk, base_step, bands, &nb, cblks, n);
    float* planes = planes_out ? planes_out : (float*)malloc(sizeof(float) * 3 * (size_t)W * H);
    or_xyz_planar_to_ict(xyz, W, H, planes);
    frame_ctx c;
    memset(&c, 0, sizeof c);
    c.planes = planes;
    c.W = W;
    c.H = H;
    c.bands = bands;
    c.cblks = cblks;
    c.thr = slope_thresh;
    c.w = W;
    c.h = H;
    for (int l = 0; l < (fourk ? 6 : 5); ++l) {
        if (c.h >= 2) par_for(3 * c.w, threads, job_cols, &c);
        if (c.w >= 2) par_for(3 * c.h, threads, job_rows, &c);
        c.w = (c.w + 1) / 2;
        c.h = (c.h + 1) / 2;
    }
    c.coeffs = coeffs ? coeffs : (uint32_t*)malloc(sizeof(uint32_t) * 1024 * (size_t)n);
    c.res = (or_t1_result*)malloc(sizeof(or_t1_result) * (size_t)n);
    c.bufs = (uint8_t*)malloc((size_t)BLOCK_BUF * n);
    par_for(n, threads, job_block, &c);
    int64_t total = 0;
    uint32_t pb = 0;
    for (int i = 0; i < n && !c.failed; ++i) {
        const or_t1_result* r = &c.res[i];
        if ((uint64_t)total + r->len > data_cap) {
            c.failed = 1;
            break;
        }
        info[i].nbps = (uint32_t)r->nbps;
        info[i].num_passes = (uint32_t)r->num_passes;
        info[i].data_offset = (uint32_t)total;
        info[i].data_len = r->len;
        info[i].pass_base = pb;
        memcpy(data + total, c.bufs + (size_t)i * BLOCK_BUF, r->len);
        for (int k = 0; k < r->num_passes; ++k) {
            pass_len[pb + k] = r->pass_len[k];
            pass_dist[pb + k] = r->pass_dist[k];
        }
        pb += (uint32_t)r->num_passes;
        total += r->len;
    }
    if (!coeffs) free(c.coeffs);
    if (!planes_out) free(planes);
    free(c.res);
    free(c.bufs);
    free(cblks);
    return c.failed ? -1 : total;
}

/* ======================================================================== decoder frame */
/* Inverse 1-D transform in float (idwtRow / idwtColumns, dwt_cpu.cpp:96-135): the low
 * half times K onto the even samples, the high half times 1/K onto the odd ones, then the
 * lifting steps delta, gamma, beta, alpha with negated coefficients. */
static void inv1d_f(float* x, int n, int stride, float* tmp) {
    if (n < 2) return;
    const int sn = (n + 1) / 2;
    for (int i = 0; i < sn; ++i) tmp[2 * i] = x[(size_t)i * stride] * (float)L_K;
    for (int i = 0; i < n / 2; ++i) tmp[2 * i + 1] = x[(size_t)(sn + i) * stride] * (float)(1.0 / L_K);
    lift_f(tmp, n, 1, 0, (float)-L_DELTA);
    lift_f(tmp, n, 1, 1, (float)-L_GAMMA);
    lift_f(tmp, n, 1, 0, (float)-L_BETA);
    lift_f(tmp, n, 1, 1, (float)-L_ALPHA);
    for (int i = 0; i < n; ++i) x[(size_t)i * stride] = tmp[i];
}

/* idwt97_2d (dwt_cpu.cpp:213-240): levels down to skip + 1, each rows then columns of the
 * level's region. */
void or_idwt97_2d_f32(float* const* planes, int nplanes, int width, int height, int levels, int skip) {
    int ws[34], hs[34];
    ws[0] = width;
    hs[0] = height;
    for (int l = 1; l <= levels; ++l) {
        ws[l] = (ws[l - 1] + 1) / 2;
        hs[l] = (hs[l - 1] + 1) / 2;
    }
    float* tmp = (float*)malloc(sizeof(float) * (size_t)imax(width, height));
    for (int l = levels; l > skip; --l) {
        const int w = ws[l - 1], h = hs[l - 1];
        for (int c = 0; c < nplanes; ++c) {
            for (int y = 0; y < h; ++y) inv1d_f(planes[c] + (size_t)y * width, w, 1, tmp);
            for (int x = 0; x < w; ++x) inv1d_f(planes[c] + x, h, width, tmp);
        }
    }
    free(tmp);
}

typedef struct {
    const uint8_t* data;
    size_t size;
    const or_cblk_code* blocks;
    const or_band* bands;
    const or_cblk* cblks;
    const int* wanted;
    float* planes[3];
    int W;
} dec_ctx;

/* One code-block of cpu_decode_backend.cpp:46-68: Tier-1 words, then each word's value
 * times half the band's step, negated for a set sign bit, into the block's place in the
 * wavelet plane. A block with no passes, or whose bytes lie past the codestream, is zero. */
static void job_decode_block(void* p, int k) {
    dec_ctx* c = (dec_ctx*)p;
    const int i = c->wanted[k];
    const or_cblk* cb = &c->cblks[i];
    const or_band* b = &c->bands[cb->band];
    const or_cblk_code* code = &c->blocks[i];
    float* plane = c->planes[b->comp] + (size_t)(b->py0 + cb->y0) * c->W + b->px0 + cb->x0;
    if (!code->passes || (size_t)code->offset + code->len > c->size) {
        for (int y = 0; y < cb->h; ++y) memset(plane + (size_t)y * c->W, 0, sizeof(float) * (size_t)cb->w);
        return;
    }
    uint32_t q[1024];
    or_t1_decode_words(c->data + code->offset, code->len, cb->w, cb->h, b->orient, code->nbps, code->passes, q);
    const float half = 0.5f * b->delta;
    for (int y = 0; y < cb->h; ++y)
        for (int x = 0; x < cb->w; ++x) {
            const uint32_t v = q[y * cb->w + x];
            const float a = (float)(v & 0x7FFFFFFFu) * half;
            plane[(size_t)y * c->W + x] = (v & 0x80000000u) ? -a : a;
        }
}

int or_decode_frame(int width, int height, int fourk, double base_step, const uint8_t* data, size_t size,
                    const or_cblk_code* blocks, int nblocks, int reduce, int high_bits, int threads, uint16_t* dst,
                    ptrdiff_t stride_bytes, float* wavelet) {
    or_band bands[OR_MAX_BANDS];
    int nb = 0;
    const int n = or_layout_make(width, height, fourk, base_step, bands, &nb, NULL, 0);
    const int levels = fourk ? 6 : 5;
    if (n < 0 || n != nblocks || reduce < 0 || reduce > levels) return -1;
    or_cblk* cblks = (or_cblk*)malloc(sizeof(or_cblk) * (size_t)n);
    or_layout_make(width, height, fourk, base_step, bands, &nb, cblks, n);
    const size_t npix = (size_t)width * height;
    float* planes = (float*)calloc(3 * npix, sizeof(float));
    int* wanted = (int*)malloc(sizeof(int) * (size_t)n);
    int nw = 0;
    for (int i = 0; i < n; ++i)
        if (bands[cblks[i].band].res <= levels - reduce) wanted[nw++] = i;
    dec_ctx c = {data, size, blocks, bands, cblks, wanted, {planes, planes + npix, planes + 2 * npix}, width};
    par_for(nw, threads, job_decode_block, &c);
    if (wavelet) memcpy(wavelet, planes, sizeof(float) * 3 * npix);
    or_idwt97_2d_f32(c.planes, 3, width, height, levels, reduce);
    /* inverse ICT (G.3), level shift, rounding and clamping (cpu_decode_backend.cpp:78-96) */
    int ow = width, oh = height;
    for (int r = 0; r < reduce; ++r) {
        ow = (ow + 1) / 2;
        oh = (oh + 1) / 2;
    }
    const int shift = high_bits ? 4 : 0;
    for (int y = 0; y < oh; ++y) {
        uint16_t* out = (uint16_t*)((uint8_t*)dst + (ptrdiff_t)y * stride_bytes);
        for (int x = 0; x < ow; ++x) {
            const size_t i = (size_t)y * width + x;
            const float y0 = c.planes[0][i], y1 = c.planes[1][i], y2 = c.planes[2][i];
            float v[3];
            v[0] = y0 + (float)ICT_INV_COLS[2][0] * y2;
            v[1] = y0 + (float)ICT_INV_COLS[1][1] * y1 + (float)ICT_INV_COLS[2][1] * y2;
            v[2] = y0 + (float)ICT_INV_COLS[1][2] * y1;
            for (int k = 0; k < 3; ++k) {
                int code = (int)floorf(v[k] + 2048.5f);
                code = code < 0 ? 0 : (code > 4095 ? 4095 : code);
                out[3 * x + k] = (uint16_t)(code << shift);
            }
        }
    }
    free(planes);
    free(wanted);
    free(cblks);
    return 0;
}
