// Colour conversion and the irreversible 9/7 DWT for CDNA4 (gfx950, wave64).
//
// DWT: one kernel launch per decomposition level covers every frame of a job, in one of
// three forms. LDS tiles (dwt_level_kernel): a workgroup owns a TX x TY tile of the level's
// LL region and all three components, stages it with a 4-sample mirrored halo in LDS, lifts
// the columns in registers, then the rows, and writes the de-interleaved results with
// coalesced row stores; levels 1 and deeper always run this way. Register strips
// (dwt_strip_kernel, dwt_strip0_kernel, dwt_strip0b_kernel): one wave per 120-column strip,
// the rows lifted through register pipelines and the row lifting's neighbours exchanged by
// DPP. Line-writing groups (dwt_stage0_kernel, level 0 by default): four such strips per
// workgroup stage their finished rows in LDS and write every band row of the group as whole
// 128-byte lines. The LL quarter goes to the next level's scratch plane, HL/LH/HH into the
// frame's coefficient planes (Mallat order; rows and level 0's high bands start on lines, at
// the column the level's descriptor gives). The order of operations on every sample is the
// CPU reference's (columns, then rows; lifting steps alpha..delta; low * (1/K), high * K),
// and the kernels are built with -ffp-contract=off, so results are bit-identical to it.
// Level 0 reads the frame's 12-bit X'Y'Z' codes and applies the DC shift and forward ICT
// on load, so the ICT planes never exist in HBM.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_types.h"
#include "kernels.h"

namespace gail_amd {

namespace {

// ISO/IEC 15444-1 Table F.4, rounded to float exactly as the CPU reference rounds them.
constexpr float kAlpha = (float)-1.586134342059924;
constexpr float kBeta = (float)-0.052980118572961;
constexpr float kGamma = (float)0.882911075530934;
constexpr float kDelta = (float)0.443506852043971;
constexpr float kLowScale = (float)(1.0 / 1.230174104914001);
constexpr float kHighScale = (float)1.230174104914001;
// Forward ICT (G.3).
constexpr float kY0[3] = {(float)0.299, (float)0.587, (float)0.114};
constexpr float kY1[3] = {(float)-0.16875, (float)-0.33126, (float)0.5};
constexpr float kY2[3] = {(float)0.5, (float)-0.41869, (float)-0.08131};

// Whole-sample symmetric extension of index i into [0, n).
__device__ __forceinline__ int mirror(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (n == 1) return 0;
    // one reflection covers every position within n - 1 of the signal
    const int r = i < 0 ? -i : 2 * (n - 1) - i;
    if ((unsigned)r < (unsigned)n) return r;
    const int period = 2 * (n - 1);
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

// Lifting of N samples held in registers, sample k at parity k & 1 (the tile origin is
// even). The first and last samples have no neighbour on one side and are left alone;
// the 4-sample halo keeps the damage outside the outputs.
template <int N> __device__ __forceinline__ void lift97(float (&v)[N]) {
#pragma unroll
    for (int k = 1; k + 1 < N; k += 2) v[k] += kAlpha * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 2; k + 1 < N; k += 2) v[k] += kBeta * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 1; k + 1 < N; k += 2) v[k] += kGamma * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 2; k + 1 < N; k += 2) v[k] += kDelta * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] *= (k & 1) ? kHighScale : kLowScale;
}

__device__ __forceinline__ int find_desc(const DwtLevelDesc* d, int n, int tile) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)d[mid].tile_base <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool kLevel0, int TX, int TY>
__global__ void __launch_bounds__(256) dwt_level_kernel(const DwtLevelDesc* __restrict__ descs, int ndesc) {
    constexpr int RX = TX + 8, RY = TY + 8, PITCH = RX + 1;
    constexpr int NPOS = RX * RY;
    constexpr int PER = (NPOS + 255) / 256;  // staged positions per thread
    __shared__ float tile[3][RY][PITCH];
    // Workgroup b runs on XCD b % 8; each XCD takes a contiguous range of tiles, so tiles
    // that share halo rows and cache lines meet in the same L2.
    const int nb = (int)gridDim.x, q = nb >> 3, rem = nb & 7;
    const int xcd = (int)blockIdx.x & 7, j = (int)blockIdx.x >> 3;
    const int tileIdx = xcd * q + min(xcd, rem) + j;
    const int di = find_desc(descs, ndesc, tileIdx);
    const DwtLevelDesc d = descs[di];
    const int t = tileIdx - d.tile_base;
    const int x0 = (t % d.tiles_x) * TX, y0 = (t / d.tiles_x) * TY;
    const int w = d.w, h = d.h;
    const int tid = threadIdx.x;
    // interior tiles need no symmetric extension
    const bool inside = x0 >= 4 && y0 >= 4 && x0 + TX + 4 <= w && y0 + TY + 4 <= h;

    // Stage the tile and its mirrored halo: every load of the thread is issued before the
    // first one is used, so they overlap.
    uint32_t at[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = tid + 256 * j;
        const int ry = i / RX, rx = i - ry * RX;
        const int gy = inside ? y0 - 4 + ry : mirror(y0 - 4 + ry, h);
        const int gx = inside ? x0 - 4 + rx : mirror(x0 - 4 + rx, w);
        at[j] = (uint32_t)gy * d.src_stride + (uint32_t)gx;
    }
    if constexpr (kLevel0) {
        const uint16_t* s0 = (const uint16_t*)d.src;
        const uint16_t* s1 = s0 + d.src_plane;
        const uint16_t* s2 = s1 + d.src_plane;
        uint32_t cx[PER], cy[PER], cz[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (NPOS % 256 == 0 || tid + 256 * j < NPOS) {
                cx[j] = s0[at[j]];
                cy[j] = s1[at[j]];
                cz[j] = s2[at[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = tid + 256 * j;
            if (NPOS % 256 == 0 || i < NPOS) {
                const int ry = i / RX, rx = i - ry * RX;
                // 12-bit codes (words above 4095 read as 4095, as the host copy clips them)
                const float x = (float)((int)min(cx[j], 4095u) - 2048);
                const float y = (float)((int)min(cy[j], 4095u) - 2048);
                const float z = (float)((int)min(cz[j], 4095u) - 2048);
                tile[0][ry][rx] = kY0[0] * x + kY0[1] * y + kY0[2] * z;
                tile[1][ry][rx] = kY1[0] * x + kY1[1] * y + kY1[2] * z;
                tile[2][ry][rx] = kY2[0] * x + kY2[1] * y + kY2[2] * z;
            }
        }
    } else {
        const float* s0 = (const float*)d.src;
        const float* s1 = s0 + d.src_plane;
        const float* s2 = s1 + d.src_plane;
        float vx[PER], vy[PER], vz[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (NPOS % 256 == 0 || tid + 256 * j < NPOS) {
                vx[j] = s0[at[j]];
                vy[j] = s1[at[j]];
                vz[j] = s2[at[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = tid + 256 * j;
            if (NPOS % 256 == 0 || i < NPOS) {
                const int ry = i / RX, rx = i - ry * RX;
                tile[0][ry][rx] = vx[j];
                tile[1][ry][rx] = vy[j];
                tile[2][ry][rx] = vz[j];
            }
        }
    }
    __syncthreads();
    // Columns (vertical lifting), one thread per column of each component.
    if (h >= 2) {
        for (int i = tid; i < 3 * RX; i += 256) {
            const int c = i / RX, rx = i - c * RX;
            float v[RY];
#pragma unroll
            for (int k = 0; k < RY; ++k) v[k] = tile[c][k][rx];
            lift97(v);
#pragma unroll
            for (int k = 4; k < TY + 4; ++k) tile[c][k][rx] = v[k];
        }
        __syncthreads();
    }
    // Rows (horizontal lifting): SPLIT threads per output row of each component, each lifting
    // an even-aligned piece of the row with its own 4-sample halo (an output depends on the 4
    // samples either side only, so the pieces compute exactly what the whole row would). All
    // pieces of a row are in the same round, and every read of a round precedes its writes.
    // Short tiles split rows in four so that most of the 256 threads have a piece.
    if (w >= 2) {
        constexpr int SPLIT = 3 * TY * 4 <= 256 ? 4 : 2;
        constexpr int PW = TX / SPLIT, HX = PW + 8;
        static_assert(PW % 2 == 0 && PW >= 8, "row pieces start at even samples");
        for (int r0 = 0; r0 < 3 * SPLIT * TY; r0 += 256) {
            const int i = r0 + tid;
            const bool act = i < 3 * SPLIT * TY;
            const int c = i / (SPLIT * TY), rr = i - c * (SPLIT * TY), ry = 4 + rr / SPLIT, xs = (rr % SPLIT) * PW;
            float v[HX];
            if (act) {
#pragma unroll
                for (int k = 0; k < HX; ++k) v[k] = tile[c][ry][xs + k];
            }
            __syncthreads();
            if (act) {
                lift97(v);
#pragma unroll
                for (int k = 4; k < PW + 4; ++k) tile[c][ry][xs + k] = v[k];
            }
            __syncthreads();
        }
    }
    // De-interleaved stores: per tile row, the even samples (low) and the odd samples
    // (high) each form a contiguous run of TX / 2 outputs. Thread tid always writes output
    // k of segment seg; the 256 threads cover 256 / TX rows per step.
    const int sny = (h + 1) >> 1;
    constexpr int HALF = TX / 2, ROWS = 256 / TX, STEPS = 3 * TY / ROWS;
    const int k = tid % HALF, seg = (tid / HALF) & 1, r0 = tid / TX;
    const int gx = x0 + 2 * k + seg;
    const int ox = seg ? (int)d.hi_off + (gx >> 1) : (gx >> 1);
    const bool colOk = gx < w;
#pragma unroll 4
    for (int j = 0; j < STEPS; ++j) {
        const int c = j / (TY / ROWS);
        const int r = r0 + ROWS * (j % (TY / ROWS));
        const int gy = y0 + r;
        if (!colOk || gy >= h) continue;
        const float v = tile[c][4 + r][4 + 2 * k + seg];
        const int oy = (gy & 1) ? sny + (gy >> 1) : (gy >> 1);
        if (!seg && !(gy & 1)) d.ll_dst[(size_t)c * d.ll_plane + (size_t)oy * d.ll_stride + ox] = v;
        else d.coef[(size_t)c * d.coef_plane + (size_t)oy * d.coef_stride + ox] = v;
    }
}

// ---------------------------------------------------------------- DWT, streaming strips
// One wave per strip of kStripOut output columns and TY output rows, all three components,
// with no LDS: lane l holds the even / odd column pair starting at X0 - 4 + 2l, so the wave
// covers the strip plus a 4-column halo each side. It streams the rows down from Y0 - 4 to
// Y0 + TY + 2 (mirrored at the frame's edges), two at a time, each column lifting in
// registers through a four-value pipeline (below); every pair of rows it finishes is lifted
// across the lanes, the neighbours' values moved one lane by DPP, and stored de-interleaved
// with coalesced row stores. The work per output sample is the lifting itself plus the
// 128 / 120 and (TY + 7) / TY halos, against 1.5 x 1.5 for the LDS tiles above.
//
// Column pipeline. After the rows up to even row k, a column holds E0 = x[k] (raw), O1 =
// x1[k-1] (odd, after the alpha step), E2 = x2[k-2] (even, after beta), O3 = x3[k-3] (odd,
// after gamma). Rows k+1 and k+2 advance it: x1[k+1], then x2[k], x3[k-1] and x4[k-2] --
// each sample's update is v + c * (left + right), in the steps' order, as in the whole-column
// lifting -- leaving even row k-2 and odd row k-1 final. The values that depend on rows
// above the first one loaded are wrong, and only rows above Y0 get them.
constexpr int kStripOut = 120;

__device__ __forceinline__ float lane_from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));  // wave_shl:1
}
__device__ __forceinline__ float lane_from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));  // wave_shr:1
}

// Horizontal lifting of one row held as (even, odd) column pairs across the wave's lanes,
// then the low / high scaling. Lanes at the wave's ends (the halo) come out wrong.
__device__ __forceinline__ void lift_row_lanes(float& e, float& o) {
    o += kAlpha * (e + lane_from_right(e));
    e += kBeta * (lane_from_left(o) + o);
    o += kGamma * (e + lane_from_right(e));
    e += kDelta * (lane_from_left(o) + o);
    e *= kLowScale;
    o *= kHighScale;
}

struct StripColumn {
    float E0, O1, E2, O3;
    // Advances by rows k+1 (odd, ro) and k+2 (even, re); returns even row k-2 and odd row
    // k-1, final and scaled.
    __device__ __forceinline__ void step(float ro, float re, float& outEven, float& outOdd) {
        const float O1n = ro + kAlpha * (E0 + re);
        const float E2n = E0 + kBeta * (O1 + O1n);
        const float O3n = O1 + kGamma * (E2 + E2n);
        const float E4 = E2 + kDelta * (O3 + O3n);
        outEven = E4 * kLowScale;
        outOdd = O3n * kHighScale;
        E0 = re;
        O1 = O1n;
        E2 = E2n;
        O3 = O3n;
    }
};

template <bool kLevel0> struct StripLoader {
    const void* src;
    uint32_t stride, plane;
    int w, h, ce;        // the lane's even column (may be outside [0, w): mirrored)
    bool inside, pairs;  // columns need no mirroring / a pair is one aligned load
    int cx0, cx1;        // mirrored columns
    __device__ __forceinline__ void init(const DwtLevelDesc& d, int x0, int lane) {
        src = d.src;
        stride = d.src_stride;
        plane = d.src_plane;
        w = d.w;
        h = d.h;
        ce = x0 - 4 + 2 * lane;
        inside = x0 >= 4 && x0 + kStripOut + 4 <= w;
        pairs = inside && (stride & 1u) == 0;
        cx0 = mirror(ce, w);
        cx1 = mirror(ce + 1, w);
    }
    // Samples (even column, odd column) of row y (mirrored) of every component; level 0
    // also applies the DC shift and the forward ICT.
    __device__ __forceinline__ void load(int y, float (&e)[3], float (&o)[3]) const {
        const uint32_t row = (uint32_t)mirror(y, h) * stride;
        if constexpr (kLevel0) {
            const uint16_t* s = (const uint16_t*)src;
            uint32_t a[3], b[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const uint16_t* r = s + (size_t)c * plane + row;
                if (pairs) {
                    const uint32_t v = *(const uint32_t*)(r + ce);
                    a[c] = v & 0xFFFFu;
                    b[c] = v >> 16;
                } else {
                    a[c] = r[cx0];
                    b[c] = r[cx1];
                }
            }
            // 12-bit codes (words above 4095 read as 4095, as the host copy clips them)
            const float x0 = (float)((int)min(a[0], 4095u) - 2048), y0 = (float)((int)min(a[1], 4095u) - 2048),
                        z0 = (float)((int)min(a[2], 4095u) - 2048);
            const float x1 = (float)((int)min(b[0], 4095u) - 2048), y1 = (float)((int)min(b[1], 4095u) - 2048),
                        z1 = (float)((int)min(b[2], 4095u) - 2048);
            e[0] = kY0[0] * x0 + kY0[1] * y0 + kY0[2] * z0;
            e[1] = kY1[0] * x0 + kY1[1] * y0 + kY1[2] * z0;
            e[2] = kY2[0] * x0 + kY2[1] * y0 + kY2[2] * z0;
            o[0] = kY0[0] * x1 + kY0[1] * y1 + kY0[2] * z1;
            o[1] = kY1[0] * x1 + kY1[1] * y1 + kY1[2] * z1;
            o[2] = kY2[0] * x1 + kY2[1] * y1 + kY2[2] * z1;
        } else {
            const float* s = (const float*)src;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float* r = s + (size_t)c * plane + row;
                if (pairs) {
                    const float2 v = *(const float2*)(r + ce);
                    e[c] = v.x;
                    o[c] = v.y;
                } else {
                    e[c] = r[cx0];
                    o[c] = r[cx1];
                }
            }
        }
    }
};

template <bool kLevel0, int TY>
__global__ void __launch_bounds__(256) dwt_strip_kernel(const DwtLevelDesc* __restrict__ descs, int ndesc, int ntiles) {
    static_assert(TY % 2 == 0, "strips start at even rows");
    // Workgroup b runs on XCD b % 8; each XCD takes a contiguous range of workgroups, so
    // strips that share halo rows and columns meet in the same L2.
    const int nb = (int)gridDim.x, q = nb >> 3, rem = nb & 7;
    const int xcd = (int)blockIdx.x & 7, j = (int)blockIdx.x >> 3;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int tileIdx = (xcd * q + min(xcd, rem) + j) * 4 + wave;
    if (tileIdx >= ntiles) return;
    const int lane = (int)threadIdx.x & 63;
    const int di = find_desc(descs, ndesc, tileIdx);
    const DwtLevelDesc d = descs[di];
    const int t = tileIdx - (int)d.tile_base;
    const int x0 = (t % d.tiles_x) * kStripOut, y0 = (t / d.tiles_x) * TY;
    const int w = d.w, h = d.h;
    StripLoader<kLevel0> ld;
    ld.init(d, x0, lane);

    // outputs: the lane's low and high columns of this level, where they are in the strip
    const int sny = (h + 1) >> 1;
    const int xo = (x0 >> 1) - 2 + lane;               // low column; the high one is hi_off + xo
    const bool inStrip = lane >= 2 && lane < 2 + kStripOut / 2;
    const bool lowOk = inStrip && ld.ce < w, highOk = inStrip && ld.ce + 1 < w;

    StripColumn col[3][2];
    {
        float e[3], o[3];
        ld.load(y0 - 4, e, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0] = StripColumn{e[c], 0.f, 0.f, 0.f};
            col[c][1] = StripColumn{o[c], 0.f, 0.f, 0.f};
        }
    }
    // steps k = y0 - 4, y0 - 2, ...: the step at k loads rows k + 1, k + 2 and finishes rows
    // k - 2 and k - 1, the strip's from k = y0 + 2 on
    float nextOddE[3], nextOddO[3], nextEvenE[3], nextEvenO[3];
    ld.load(y0 - 3, nextOddE, nextOddO);
    ld.load(y0 - 2, nextEvenE, nextEvenO);
    for (int k = y0 - 4; k <= y0 + TY; k += 2) {
        float oddE[3], oddO[3], evenE[3], evenO[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            oddE[c] = nextOddE[c];
            oddO[c] = nextOddO[c];
            evenE[c] = nextEvenE[c];
            evenO[c] = nextEvenO[c];
        }
        if (k + 2 <= y0 + TY) {  // the next step's rows, in flight while this one computes
            ld.load(k + 3, nextOddE, nextOddO);
            ld.load(k + 4, nextEvenE, nextEvenO);
        }
        float lowE[3], lowO[3], highE[3], highO[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0].step(oddE[c], evenE[c], lowE[c], highE[c]);
            col[c][1].step(oddO[c], evenO[c], lowO[c], highO[c]);
        }
        const int gy = k - 2;  // the even row finished; gy + 1 the odd one
        if (gy < y0 || gy >= h) continue;  // wave-uniform
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lift_row_lanes(lowE[c], lowO[c]);
            lift_row_lanes(highE[c], highO[c]);
            const int oyL = gy >> 1;
            if (lowOk) d.ll_dst[(size_t)c * d.ll_plane + (size_t)oyL * d.ll_stride + xo] = lowE[c];
            if (highOk) d.coef[(size_t)c * d.coef_plane + (size_t)oyL * d.coef_stride + d.hi_off + xo] = lowO[c];
            if (gy + 1 < h) {
                const int oyH = sny + (gy >> 1);
                if (lowOk) d.coef[(size_t)c * d.coef_plane + (size_t)oyH * d.coef_stride + xo] = highE[c];
                if (highOk) d.coef[(size_t)c * d.coef_plane + (size_t)oyH * d.coef_stride + d.hi_off + xo] = highO[c];
            }
        }
    }
}

// Level 0 as strips of kStripOut x TY outputs with PRE input rows loaded ahead: a lane issues
// the loads of its column pair (the 12-bit codes of both columns in one 32-bit word per
// component) for the first PRE rows before it converts or lifts anything, and keeps PRE rows
// in flight from then on, so a wave has much of its input on the way at once, and short
// strips give a 4K batch several times as many waves as the streaming strips. The lifting is
// the streaming strips' own, operation for operation: the ICT on each row as it enters the
// column pipelines, the same pipeline steps, the same row lifting across lanes.
// Global-memory views of the frame buffers (the descriptors hold generic pointers; through
// these the compiler issues global rather than flat loads and stores, which the scalar-memory
// wait counter does not track).
typedef __attribute__((address_space(1))) const uint16_t GlobalU16;
typedef __attribute__((address_space(1))) const uint32_t GlobalU32;
typedef __attribute__((address_space(1))) float GlobalF32;
__device__ __forceinline__ void store_global(float* p, float v) { *(GlobalF32*)p = v; }

// One strip. kInner: every loaded row and column lies inside the frame and its rows are
// even-aligned in memory (pairs of columns in one load, rows without mirroring, every output
// column present), so the inner strips of a frame run without any per-row or per-store test.
template <int TY, int PRE, bool kInner>
__device__ __forceinline__ void strip0_body(const DwtLevelDesc& d, int x0, int y0, int lane) {
    constexpr int RY = TY + 7;  // rows y0 - 4 .. y0 + TY + 2
    const int w = d.w, h = d.h;
    const int ce = x0 - 4 + 2 * lane;
    const int cx0 = kInner ? ce : mirror(ce, w), cx1 = kInner ? ce + 1 : mirror(ce + 1, w);

    // the codes of both columns of every row, each component's pair in one word; the row's
    // start is the same for the whole wave, the lane's column an offset from it
    const uint16_t* const s0 = (const uint16_t*)d.src;
    uint32_t code[RY][3];
    auto loadRow = [&](int r) {
        const int y = kInner ? y0 - 4 + r : mirror(y0 - 4 + r, h);
        const size_t row = (size_t)y * d.src_stride;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            GlobalU16* const rowp = (GlobalU16*)(s0 + (size_t)c * d.src_plane + row);
            if constexpr (kInner)
                code[r][c] = ((GlobalU32*)rowp)[(uint32_t)ce >> 1];
            else
                code[r][c] = (uint32_t)rowp[(uint32_t)cx0] | (uint32_t)rowp[(uint32_t)cx1] << 16;
        }
    };
#pragma unroll
    for (int r = 0; r < (PRE < RY ? PRE : RY); ++r) loadRow(r);
    // DC shift and forward ICT of row r, as StripLoader<true>::load computes them
    auto rowSamples = [&](int r, float (&e)[3], float (&o)[3]) {
        // 12-bit codes (words above 4095 read as 4095, as the host copy clips them)
        const float x0f = (float)((int)min(code[r][0] & 0xFFFFu, 4095u) - 2048);
        const float y0f = (float)((int)min(code[r][1] & 0xFFFFu, 4095u) - 2048);
        const float z0f = (float)((int)min(code[r][2] & 0xFFFFu, 4095u) - 2048);
        const float x1f = (float)((int)min(code[r][0] >> 16, 4095u) - 2048);
        const float y1f = (float)((int)min(code[r][1] >> 16, 4095u) - 2048);
        const float z1f = (float)((int)min(code[r][2] >> 16, 4095u) - 2048);
        e[0] = kY0[0] * x0f + kY0[1] * y0f + kY0[2] * z0f;
        e[1] = kY1[0] * x0f + kY1[1] * y0f + kY1[2] * z0f;
        e[2] = kY2[0] * x0f + kY2[1] * y0f + kY2[2] * z0f;
        o[0] = kY0[0] * x1f + kY0[1] * y1f + kY0[2] * z1f;
        o[1] = kY1[0] * x1f + kY1[1] * y1f + kY1[2] * z1f;
        o[2] = kY2[0] * x1f + kY2[1] * y1f + kY2[2] * z1f;
    };

    const int sny = (h + 1) >> 1;
    const int xo = (x0 >> 1) - 2 + lane;
    const bool inStrip = lane >= 2 && lane < 2 + kStripOut / 2;
    const bool lowOk = inStrip && (kInner || ce < w), highOk = inStrip && (kInner || ce + 1 < w);

    StripColumn col[3][2];
    {
        float e[3], o[3];
        rowSamples(0, e, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0] = StripColumn{e[c], 0.f, 0.f, 0.f};
            col[c][1] = StripColumn{o[c], 0.f, 0.f, 0.f};
        }
    }
    // the step at k = y0 - 4 + 2 i takes rows k + 1, k + 2 (preloaded rows 2 i + 1, 2 i + 2)
    // and finishes rows k - 2, k - 1
#pragma unroll
    for (int i = 0; i <= (TY + 4) / 2; ++i) {
        const int k = y0 - 4 + 2 * i;
        if (PRE + 2 * i < RY) loadRow(PRE + 2 * i);  // keep PRE rows ahead in flight
        if (PRE + 2 * i + 1 < RY) loadRow(PRE + 2 * i + 1);
        float oddE[3], oddO[3], evenE[3], evenO[3];
        rowSamples(2 * i + 1, oddE, oddO);
        rowSamples(2 * i + 2, evenE, evenO);
        float lowE[3], lowO[3], highE[3], highO[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0].step(oddE[c], evenE[c], lowE[c], highE[c]);
            col[c][1].step(oddO[c], evenO[c], lowO[c], highO[c]);
        }
        const int gy = k - 2;
        if (2 * i < 6) continue;                 // the rows above the strip: gy < y0
        if (!kInner && gy >= h) continue;        // wave-uniform
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lift_row_lanes(lowE[c], lowO[c]);
            lift_row_lanes(highE[c], highO[c]);
        }
        // the rows' starts are the same for the whole wave, the lane's column an offset
        // (unsigned: lanes left of the frame store nothing)
        const uint32_t xu = (uint32_t)xo;
        const int oyL = gy >> 1, oyH = sny + (gy >> 1);
        const bool oddRow = kInner || gy + 1 < h;
        if (kInner) {
            if (inStrip) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float* const llRow = d.ll_dst + (size_t)c * d.ll_plane + (size_t)oyL * d.ll_stride;
                    float* const hlRow = d.coef + (size_t)c * d.coef_plane + (size_t)oyL * d.coef_stride + d.hi_off;
                    float* const lhRow = d.coef + (size_t)c * d.coef_plane + (size_t)oyH * d.coef_stride;
                    store_global(llRow + xu, lowE[c]);
                    store_global(hlRow + xu, lowO[c]);
                    store_global(lhRow + xu, highE[c]);
                    store_global(lhRow + d.hi_off + xu, highO[c]);
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float* const llRow = d.ll_dst + (size_t)c * d.ll_plane + (size_t)oyL * d.ll_stride;
                float* const hlRow = d.coef + (size_t)c * d.coef_plane + (size_t)oyL * d.coef_stride + d.hi_off;
                if (lowOk) store_global(llRow + xu, lowE[c]);
                if (highOk) store_global(hlRow + xu, lowO[c]);
                if (oddRow) {
                    float* const lhRow = d.coef + (size_t)c * d.coef_plane + (size_t)oyH * d.coef_stride;
                    if (lowOk) store_global(lhRow + xu, highE[c]);
                    if (highOk) store_global(lhRow + d.hi_off + xu, highO[c]);
                }
            }
        }
    }
}

template <int TY, int PRE, int MINW>
__global__ void __launch_bounds__(256, MINW) dwt_strip0_kernel(const DwtLevelDesc* __restrict__ descs, int ndesc, int ntiles) {
    static_assert(TY % 2 == 0, "strips start at even rows");
    static_assert(PRE >= 3, "rows 2i + 1 and 2i + 2 must be loaded by step i");
    const int nb = (int)gridDim.x, q = nb >> 3, rem = nb & 7;
    const int xcd = (int)blockIdx.x & 7, j = (int)blockIdx.x >> 3;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int tileIdx = (xcd * q + min(xcd, rem) + j) * 4 + wave;
    if (tileIdx >= ntiles) return;
    const int lane = (int)threadIdx.x & 63;
    const int di = find_desc(descs, ndesc, tileIdx);
    const DwtLevelDesc d = descs[di];
    const int t = tileIdx - (int)d.tile_base;
    const int x0 = (t % d.tiles_x) * kStripOut, y0 = (t / d.tiles_x) * TY;
    const bool inner = x0 >= 4 && x0 + kStripOut + 4 <= d.w && y0 >= 4 && y0 + TY + 3 <= d.h &&
                       ((d.src_stride | d.src_plane) & 1u) == 0;
    if (inner) strip0_body<TY, PRE, true>(d, x0, y0, lane);
    else strip0_body<TY, PRE, false>(d, x0, y0, lane);
}

// Level 0 in 120 x TY strips with the rows loaded ahead, as dwt_strip0_kernel, but every
// load and store a raw buffer access: the frame's base in a buffer resource, the row and
// component a scalar byte offset, the lane's column a fixed vector offset, so no vector
// instruction computes an address. A lane with nothing to store (outside the strip or the
// frame) stores to offset kDiscard, beyond the resource's range, which the hardware drops:
// the stores of the inner strips run with no branch and no per-step vector work. Byte
// offsets are 32-bit: a frame's three planes stay below 2 GiB (the host checks it). The
// codes' DC shift runs on both halves of a word at once. The arithmetic is otherwise the
// streaming strips' own, operation for operation.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_rsrc(const void* p) {
    // raw (stride 0) resource of 2^31 - 1 bytes; word 3 is the gfx9 default data format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7FFFFFFF, 0x00020000);
}
// A vector offset beyond every frame_rsrc range: stores to it are dropped.
constexpr uint32_t kDiscard = 0x80000000u;

// Row y of a strip mirrored into [0, h) for h >= 11: one reflection is exact for the rows
// -4 .. h + 9 a strip of up to 8 rows reads; rows further down land on some valid row.
__device__ __forceinline__ int mirror_row(int y, int h) {
    y = y < 0 ? -y : y;
    y = y < h ? y : 2 * (h - 1) - y;
    return y > 0 ? y : 0;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
// The DC-shifted samples of the two 12-bit codes of a word (low half first): both halves
// clipped to 4095 (words above it read as 4095, as the host copy clips them) and shifted by
// -2048 with one packed 16-bit operation each, then converted from the sign-extended halves.
// The values are (float)((int)min(code, 4095) - 2048) of each half.
__device__ __forceinline__ void dc_shift_pair(uint32_t code, float& lo, float& hi) {
    u16x2 v = __builtin_bit_cast(u16x2, code);
    v = __builtin_elementwise_min(v, (u16x2){4095, 4095});
    const s16x2 s = __builtin_bit_cast(s16x2, v) - (s16x2){2048, 2048};
    lo = (float)s.x;
    hi = (float)s.y;
}

template <int TY, int PRE, bool kInner>
__device__ __forceinline__ void strip0b_body(const DwtLevelDesc& d, int x0, int y0, int lane) {
    constexpr int RY = TY + 7;  // rows y0 - 4 .. y0 + TY + 2
    static_assert(TY <= 8, "mirror_row covers strips of up to 8 rows");
    const int w = d.w, h = d.h;
    const int ce = x0 - 4 + 2 * lane;
    const __amdgpu_buffer_rsrc_t src = frame_rsrc(d.src);
    // kInner: the pair (ce, ce + 1) is one aligned word; else two mirrored columns
    const uint32_t voE = 2u * (uint32_t)(kInner ? ce : mirror(ce, w));
    const uint32_t voO = kInner ? 0u : 2u * (uint32_t)mirror(ce + 1, w);
    const uint32_t sRow = 2u * d.src_stride, sPlane = 2u * d.src_plane;
    uint32_t code[RY][3];
    auto loadRow = [&](int r) {
        const int y = kInner ? y0 - 4 + r : mirror_row(y0 - 4 + r, h);
        const uint32_t so = (uint32_t)y * sRow;  // wave-uniform
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t soc = so + (uint32_t)c * sPlane;
            if constexpr (kInner)
                code[r][c] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(src, (int)voE, (int)soc, 0);
            else
                code[r][c] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(src, (int)voE, (int)soc, 0) |
                             (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(src, (int)voO, (int)soc, 0) << 16;
        }
    };
#pragma unroll
    for (int r = 0; r < (PRE < RY ? PRE : RY); ++r) loadRow(r);
    // DC shift and forward ICT of row r, as StripLoader<true>::load computes them
    auto rowSamples = [&](int r, float (&e)[3], float (&o)[3]) {
        float x0f, y0f, z0f, x1f, y1f, z1f;
        dc_shift_pair(code[r][0], x0f, x1f);
        dc_shift_pair(code[r][1], y0f, y1f);
        dc_shift_pair(code[r][2], z0f, z1f);
        e[0] = kY0[0] * x0f + kY0[1] * y0f + kY0[2] * z0f;
        e[1] = kY1[0] * x0f + kY1[1] * y0f + kY1[2] * z0f;
        e[2] = kY2[0] * x0f + kY2[1] * y0f + kY2[2] * z0f;
        o[0] = kY0[0] * x1f + kY0[1] * y1f + kY0[2] * z1f;
        o[1] = kY1[0] * x1f + kY1[1] * y1f + kY1[2] * z1f;
        o[2] = kY2[0] * x1f + kY2[1] * y1f + kY2[2] * z1f;
    };

    const __amdgpu_buffer_rsrc_t coef = frame_rsrc(d.coef), llr = frame_rsrc(d.ll_dst);
    const int sny = (h + 1) >> 1;
    const int xo = (x0 >> 1) - 2 + lane;  // the lane's low column
    const bool inStrip = lane >= 2 && lane < 2 + kStripOut / 2;
    const uint32_t voL = inStrip && (kInner || ce < w) ? 4u * (uint32_t)xo : kDiscard;
    const uint32_t voH = inStrip && (kInner || ce + 1 < w) ? 4u * (d.hi_off + (uint32_t)xo) : kDiscard;
    const uint32_t llRow = 4u * d.ll_stride, llPlane = 4u * d.ll_plane;
    const uint32_t cRow = 4u * d.coef_stride, cPlane = 4u * d.coef_plane;

    StripColumn col[3][2];
    {
        float e[3], o[3];
        rowSamples(0, e, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0] = StripColumn{e[c], 0.f, 0.f, 0.f};
            col[c][1] = StripColumn{o[c], 0.f, 0.f, 0.f};
        }
    }
    // the step at k = y0 - 4 + 2 i takes rows k + 1, k + 2 (preloaded rows 2 i + 1, 2 i + 2)
    // and finishes rows k - 2, k - 1
#pragma unroll
    for (int i = 0; i <= (TY + 4) / 2; ++i) {
        const int k = y0 - 4 + 2 * i;
        if (PRE + 2 * i < RY) loadRow(PRE + 2 * i);  // keep PRE rows ahead in flight
        if (PRE + 2 * i + 1 < RY) loadRow(PRE + 2 * i + 1);
        float oddE[3], oddO[3], evenE[3], evenO[3];
        rowSamples(2 * i + 1, oddE, oddO);
        rowSamples(2 * i + 2, evenE, evenO);
        float lowE[3], lowO[3], highE[3], highO[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0].step(oddE[c], evenE[c], lowE[c], highE[c]);
            col[c][1].step(oddO[c], evenO[c], lowO[c], highO[c]);
        }
        const int gy = k - 2;
        if (2 * i < 6) continue;           // the rows above the strip: gy < y0
        if (!kInner && gy >= h) continue;  // wave-uniform
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lift_row_lanes(lowE[c], lowO[c]);
            lift_row_lanes(highE[c], highO[c]);
        }
        const bool oddIn = kInner || gy + 1 < h;
        const uint32_t vLo = oddIn ? voL : kDiscard, vHo = oddIn ? voH : kDiscard;
        const uint32_t oyL = (uint32_t)gy >> 1, oyH = (uint32_t)sny + oyL;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t soLL = (uint32_t)c * llPlane + oyL * llRow;
            const uint32_t soHL = (uint32_t)c * cPlane + oyL * cRow;
            const uint32_t soLH = (uint32_t)c * cPlane + oyH * cRow;
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(lowE[c]), llr, (int)voL, (int)soLL, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(lowO[c]), coef, (int)voH, (int)soHL, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(highE[c]), coef, (int)vLo, (int)soLH, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(highO[c]), coef, (int)vHo, (int)soLH, 0);
        }
    }
}

// WG waves per workgroup, one strip each.
template <int TY, int PRE, int MINW, int WG>
__global__ void __launch_bounds__(64 * WG, MINW) dwt_strip0b_kernel(const DwtLevelDesc* __restrict__ descs, int ndesc, int ntiles) {
    static_assert(TY % 2 == 0, "strips start at even rows");
    static_assert(PRE >= 3, "rows 2i + 1 and 2i + 2 must be loaded by step i");
    const int nb = (int)gridDim.x, q = nb >> 3, rem = nb & 7;
    const int xcd = (int)blockIdx.x & 7, j = (int)blockIdx.x >> 3;
    const int wave = WG == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int tileIdx = (xcd * q + min(xcd, rem) + j) * WG + wave;
    if (tileIdx >= ntiles) return;
    const int lane = (int)threadIdx.x & 63;
    const int di = find_desc(descs, ndesc, tileIdx);
    const DwtLevelDesc d = descs[di];
    const int t = tileIdx - (int)d.tile_base;
    const int x0 = (t % d.tiles_x) * kStripOut, y0 = (t / d.tiles_x) * TY;
    const bool inner = x0 >= 4 && x0 + kStripOut + 4 <= d.w && y0 >= 4 && y0 + TY + 3 <= d.h &&
                       ((d.src_stride | d.src_plane) & 1u) == 0;
    if (inner) strip0b_body<TY, PRE, true>(d, x0, y0, lane);
    else strip0b_body<TY, PRE, false>(d, x0, y0, lane);
}

// Level 0 in workgroups of NW adjacent 120 x 8 strips, one per wave with its rows loaded
// ahead as in dwt_strip0b_kernel, whose outputs leave the CU as whole 128-byte lines: each
// pair of finished rows is staged in LDS, then the workgroup writes the group's run of
// NW * 60 columns of every band row with 16-byte stores. Coefficient rows, level 0's LL rows
// and its high bands all start on a line (the padded coefficient layout) and a group's first
// output column is 60 NW times its index (for four strips a multiple of 16 floats), so a run
// starts on a line or half-way into one and covers whole lines from there; a strip's own
// 240-byte pieces would start anywhere in a line, and HBM takes partial lines far more slowly. The
// staging is double-buffered: the pair finished at step i goes to buffer (i - 3) & 1, so one
// barrier per pair orders the writes of a buffer after the reads of its previous pair.
template <int NW, int MINW>
__global__ void __launch_bounds__(64 * NW, MINW) dwt_stage0_kernel(const DwtLevelDesc* __restrict__ descs, int ndesc, int ngroups) {
    constexpr int TY = 8, RY = TY + 7, RUN = NW * (kStripOut / 2), Q = RUN / 4;
    static_assert(RUN % 8 == 0 && 12 % NW == 0 && Q <= 64, "runs start on 32-byte boundaries, whole runs per wave");
    __shared__ float4 stage[2][12][Q];  // buffer, run (component c, band b: c + 3 b), quads
    const int nb = (int)gridDim.x, q = nb >> 3, rem = nb & 7;
    const int xcd = (int)blockIdx.x & 7, j = (int)blockIdx.x >> 3;
    const int g = xcd * q + min(xcd, rem) + j;
    if (g >= ngroups) return;  // the whole workgroup
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    const int di = find_desc(descs, ndesc, g);
    const DwtLevelDesc d = descs[di];
    const int t = g - (int)d.tile_base;
    const int gx = t % d.tiles_x, y0 = (t / d.tiles_x) * TY;
    const int x0 = gx * (NW * kStripOut) + wave * kStripOut;
    const int w = d.w, h = d.h;
    const bool act = x0 < w;  // waves right of the frame only keep the barriers
    const bool inner = act && x0 >= 4 && x0 + kStripOut + 4 <= w && y0 >= 4 && y0 + TY + 3 <= h &&
                       ((d.src_stride | d.src_plane) & 1u) == 0;
    const int ce = x0 - 4 + 2 * lane;

    uint32_t code[RY][3];
    {
        const __amdgpu_buffer_rsrc_t src = frame_rsrc(d.src);
        const uint32_t sRow = 2u * d.src_stride, sPlane = 2u * d.src_plane;
        if (inner) {
            const uint32_t voE = 2u * (uint32_t)ce;
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const uint32_t so = (uint32_t)(y0 - 4 + r) * sRow;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    code[r][c] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(src, (int)voE, (int)(so + (uint32_t)c * sPlane), 0);
            }
        } else if (act) {
            const uint32_t voE = 2u * (uint32_t)mirror(ce, w), voO = 2u * (uint32_t)mirror(ce + 1, w);
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const uint32_t so = (uint32_t)mirror_row(y0 - 4 + r, h) * sRow;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int soc = (int)(so + (uint32_t)c * sPlane);
                    code[r][c] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(src, (int)voE, soc, 0) |
                                 (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(src, (int)voO, soc, 0) << 16;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < RY; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) code[r][c] = 0;
        }
    }
    // DC shift and forward ICT of row r, as StripLoader<true>::load computes them
    auto rowSamples = [&](int r, float (&e)[3], float (&o)[3]) {
        float x0f, y0f, z0f, x1f, y1f, z1f;
        dc_shift_pair(code[r][0], x0f, x1f);
        dc_shift_pair(code[r][1], y0f, y1f);
        dc_shift_pair(code[r][2], z0f, z1f);
        e[0] = kY0[0] * x0f + kY0[1] * y0f + kY0[2] * z0f;
        e[1] = kY1[0] * x0f + kY1[1] * y0f + kY1[2] * z0f;
        e[2] = kY2[0] * x0f + kY2[1] * y0f + kY2[2] * z0f;
        o[0] = kY0[0] * x1f + kY0[1] * y1f + kY0[2] * z1f;
        o[1] = kY1[0] * x1f + kY1[1] * y1f + kY1[2] * z1f;
        o[2] = kY2[0] * x1f + kY2[1] * y1f + kY2[2] * z1f;
    };

    const bool inStrip = lane >= 2 && lane < 2 + kStripOut / 2;
    const bool lowOk = act && inStrip && ce < w, highOk = act && inStrip && ce + 1 < w;
    float* const slot = (float*)&stage[0][0][0] + wave * (kStripOut / 2) + lane - 2;  // run-relative
    const int sny = (h + 1) >> 1, xo0 = gx * RUN;
    const int nLow = min(RUN, ((w + 1) >> 1) - xo0), nHigh = min(RUN, (w >> 1) - xo0);
    const __amdgpu_buffer_rsrc_t coef = frame_rsrc(d.coef), llr = frame_rsrc(d.ll_dst);
    const uint32_t llRow = 4u * d.ll_stride, llPlane = 4u * d.ll_plane;
    const uint32_t cRow = 4u * d.coef_stride, cPlane = 4u * d.coef_plane, hiB = 4u * d.hi_off;
    const uint32_t voQ = 4u * (uint32_t)(xo0 + 4 * lane);  // this lane's quad in every run

    StripColumn col[3][2];
    {
        float e[3], o[3];
        rowSamples(0, e, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0] = StripColumn{e[c], 0.f, 0.f, 0.f};
            col[c][1] = StripColumn{o[c], 0.f, 0.f, 0.f};
        }
    }
#pragma unroll
    for (int i = 0; i <= (TY + 4) / 2; ++i) {
        float oddE[3], oddO[3], evenE[3], evenO[3];
        rowSamples(2 * i + 1, oddE, oddO);
        rowSamples(2 * i + 2, evenE, evenO);
        float lowE[3], lowO[3], highE[3], highO[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            col[c][0].step(oddE[c], evenE[c], lowE[c], highE[c]);
            col[c][1].step(oddO[c], evenO[c], lowO[c], highO[c]);
        }
        if (i < 3) continue;               // the rows above the group
        const int gy = y0 - 6 + 2 * i;     // even row finished; gy + 1 the odd one
        if (gy >= h) continue;             // the same for the whole workgroup
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lift_row_lanes(lowE[c], lowO[c]);
            lift_row_lanes(highE[c], highO[c]);
        }
        const int buf = (i - 3) & 1;
        float* const sb = slot + buf * (12 * RUN);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (lowOk) {
                sb[(c + 0) * RUN] = lowE[c];   // LL
                sb[(c + 6) * RUN] = highE[c];  // LH
            }
            if (highOk) {
                sb[(c + 3) * RUN] = lowO[c];   // HL
                sb[(c + 9) * RUN] = highO[c];  // HH
            }
        }
        __syncthreads();
        // wave v writes runs v, v + NW, ...: lane l the quad at columns 4l .. 4l + 3 of the
        // run, a 16-byte store at a fixed vector offset; the run's row, component and band
        // are the scalar offset
        const bool oddIn = gy + 1 < h;  // LH and HH only where the odd row exists
        const uint32_t oyL = (uint32_t)gy >> 1, oyH = (uint32_t)sny + oyL;
#pragma unroll
        for (int k = 0; k < 12 / NW; ++k) {
            const int run = wave + NW * k;  // wave-uniform
            const int c = run % 3, band = run / 3;
            if (band >= 2 && !oddIn) continue;
            const int n = (band & 1) ? nHigh : nLow;
            const uint32_t so = band == 0 ? (uint32_t)c * llPlane + oyL * llRow
                                          : (uint32_t)c * cPlane + (band >= 2 ? oyH : oyL) * cRow + ((band & 1) ? hiB : 0u);
            const __amdgpu_buffer_rsrc_t dst = band == 0 ? llr : coef;
            if (lane < Q && 4 * lane < n) {
                const float4 v = stage[buf][run][lane];
                if (4 * lane + 4 <= n) {
                    const u32x4 q4 = {__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)};
                    __builtin_amdgcn_raw_buffer_store_b128(q4, dst, (int)voQ, (int)so, 0);
                } else {  // the frame's last columns
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.x), dst, (int)voQ, (int)so, 0);
                    if (4 * lane + 1 < n) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.y), dst, (int)voQ + 4, (int)so, 0);
                    if (4 * lane + 2 < n) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.z), dst, (int)voQ + 8, (int)so, 0);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- colour
// Source R'G'B' sample (0..1 after range handling) to the 16-bit code the CPU reference
// rounds it to (color.cpp:136-139); the per-pixel path then works on code / 65535.
__device__ __forceinline__ uint32_t code16(float v) {
    v = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
    const float s = v * 65535.f;
    const int r = (int)s;
    return (uint32_t)(r + ((s - (float)r) >= 0.5f ? 1 : 0));
}

// EOTF table index of a source value e in 0..1 (color_math.h:140-142).
__device__ __forceinline__ int lut_index(float e) {
    return (int)floorf((e < 0.f ? 0.f : (e > 1.f ? 1.f : e)) * 65535.0f + 0.5f);
}

// powf as the host C library computes it in practice: the power in double, rounded once to
// float (the library's own double evaluation rounds to the same float except within a tiny
// fraction of an ulp of a rounding boundary).
__device__ __forceinline__ float pow_f(float x, float y) { return (float)pow((double)x, (double)y); }

// SMPTE ST 2084 in float (color_math.h:27-40).
__device__ __forceinline__ float pq_to_nits(float e) {
    const float m1 = 2610.0f / 16384.0f, m2 = 2523.0f / 4096.0f * 128.0f, c1 = 3424.0f / 4096.0f,
                c2 = 2413.0f / 4096.0f * 32.0f, c3 = 2392.0f / 4096.0f * 32.0f;
    const float p = pow_f(fminf(fmaxf(e, 0.0f), 1.0f), 1.0f / m2);
    return 10000.0f * pow_f(fmaxf(p - c1, 0.0f) / (c2 - c3 * p), 1.0f / m1);
}
__device__ __forceinline__ float nits_to_pq(float nits) {
    const float m1 = 2610.0f / 16384.0f, m2 = 2523.0f / 4096.0f * 128.0f, c1 = 3424.0f / 4096.0f,
                c2 = 2413.0f / 4096.0f * 32.0f, c3 = 2392.0f / 4096.0f * 32.0f;
    const float y = pow_f(fminf(fmaxf(nits / 10000.0f, 0.0f), 1.0f), m1);
    return pow_f((c1 + c2 * y) / (1.0f + c3 * y), m2);
}

// Tone curve: target luminance (cd/m^2) for source luminance y (color_math.h:42-65). The
// shoulder is scene-relative (linear to the knee, then an extended-Reinhard roll-off that
// reaches the peak at the source peak); BT.2390 is the EETF in the PQ domain.
__device__ __forceinline__ float tone_curve(const float* cp, float y) {
    if (y <= 0.0f) return 0.0f;
    if (cp[12] > 1.5f) {
        const float srcMax = cp[39], ks = cp[41];
        const float e1 = nits_to_pq(y) / srcMax;
        if (e1 < ks) return y;
        const float t = (fminf(e1, 1.0f) - ks) / (1.0f - ks);
        const float t2 = t * t, t3 = t2 * t;
        const float e2 = (2.0f * t3 - 3.0f * t2 + 1.0f) * ks + (t3 - 2.0f * t2 + t) * (1.0f - ks) +
                         (-2.0f * t3 + 3.0f * t2) * cp[40];
        return pq_to_nits(e2 * srcMax);
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

// Gamut compression (color_math.h:67-105): toward neutral grey of the same luminance until
// every target-gamut channel lies in [0, 1], smoothly beyond the threshold.
__device__ __forceinline__ void compress_gamut(const float* cp, float& X, float& Y, float& Z) {
    float r = cp[20] * X + cp[21] * Y + cp[22] * Z;
    float g = cp[23] * X + cp[24] * Y + cp[25] * Z;
    float b = cp[26] * X + cp[27] * Y + cp[28] * Z;
    const float n = Y / cp[17];
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
    X = cp[29] * r + cp[30] * g + cp[31] * b;
    Y = cp[32] * r + cp[33] * g + cp[34] * b;
    Z = cp[35] * r + cp[36] * g + cp[37] * b;
}

// cd/m^2 -> 12-bit code: the number of code thresholds at or below v. The thresholds are
// where the host's floor(4095 * (v / 52.37)^(1 / 2.6) + 0.5) steps up, which is monotone
// in v, so this reproduces the CPU reference's codes exactly.
__device__ __forceinline__ uint32_t encode_code(const float* __restrict__ thr, float v) {
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 2048; step; step >>= 1)
        if (lo + step <= 4095u && thr[lo + step] <= v) lo += step;
    return lo;
}

__device__ __forceinline__ float sample(const void* src, int bits, size_t i) {
    return bits > 8 ? (float)((const uint16_t*)src)[i] : (float)((const uint8_t*)src)[i];
}

}  // namespace

// Any input format -> planar 12-bit X'Y'Z' codes (one thread per container pixel).
__global__ void __launch_bounds__(256) color_kernel(ColorDesc d) {
    __shared__ float thr[4096];
    for (int i = threadIdx.x; i < 4096; i += 256) thr[i] = d.thresholds[i];
    __syncthreads();
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t npix = (uint32_t)d.width * d.height;
    if (i >= npix) return;
    const int x = (int)(i % (uint32_t)d.width), y = (int)(i / (uint32_t)d.width);
    uint32_t code[3] = {0, 0, 0};  // source R'G'B' as 16-bit codes
    if (d.kind == 0) {  // interleaved 16-bit R'G'B'
        const uint16_t* s = (const uint16_t*)d.src + 3 * (size_t)i;
        for (int c = 0; c < 3; ++c) code[c] = s[c];
    } else {  // planar Y'CbCr or G'B'R' of the active picture, black around it
        const int ax = x - d.ox, ay = y - d.oy;
        float f[3] = {0.f, 0.f, 0.f};
        if (ax >= 0 && ay >= 0 && ax < d.aw && ay < d.ah) {
            const int cw = (d.aw + d.sx) >> d.sx, ch = (d.ah + d.sy) >> d.sy;
            const size_t uoff = (size_t)d.aw * d.ah, voff = uoff + (size_t)cw * ch;
            float cyp = !d.sy ? (float)ay : d.chroma == 2 ? ay * 0.5f : (ay - 0.5f) * 0.5f;
            cyp = cyp < 0.f ? 0.f : (cyp > (float)(ch - 1) ? (float)(ch - 1) : cyp);
            float cxp = !d.sx ? (float)ax : d.chroma == 1 ? (ax - 0.5f) * 0.5f : ax * 0.5f;
            cxp = cxp < 0.f ? 0.f : (cxp > (float)(cw - 1) ? (float)(cw - 1) : cxp);
            const int yy0 = (int)cyp, yy1 = min(yy0 + 1, ch - 1), xx0 = (int)cxp, xx1 = min(xx0 + 1, cw - 1);
            const float fy = cyp - yy0, fx = cxp - xx0;
            float uv[2];
            for (int k = 0; k < 2; ++k) {
                const size_t off = k ? voff : uoff;
                const float a = sample(d.src, d.bits, off + (size_t)yy0 * cw + xx0);
                const float b = sample(d.src, d.bits, off + (size_t)yy0 * cw + xx1);
                const float c = sample(d.src, d.bits, off + (size_t)yy1 * cw + xx0);
                const float e = sample(d.src, d.bits, off + (size_t)yy1 * cw + xx1);
                uv[k] = (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + e * fx) * fy;
            }
            float Y = sample(d.src, d.bits, (size_t)ay * d.aw + ax), U = uv[0], V = uv[1];
            const float maxv = (float)((1 << d.bits) - 1), scale = (float)(1 << (d.bits - 8));
            if (d.rgb) {
                f[0] = V / maxv;
                f[1] = Y / maxv;
                f[2] = U / maxv;
            } else {
                if (d.full_range) {
                    Y /= maxv;
                    U = (U - 128.f * scale) / maxv;
                    V = (V - 128.f * scale) / maxv;
                } else {
                    Y = (Y - 16.f * scale) / (219.f * scale);
                    U = (U - 128.f * scale) / (224.f * scale);
                    V = (V - 128.f * scale) / (224.f * scale);
                }
                f[0] = Y + d.crR * V;
                f[1] = Y - d.cbG * U - d.crG * V;
                f[2] = Y + d.cbB * U;
            }
        }
        for (int c = 0; c < 3; ++c) code[c] = code16(f[c]);
    }
    // cm_rgb_to_xyz12 (color_math.h:114-179) on code / 65535
    const float e0 = (float)code[0] / 65535.f, e1 = (float)code[1] / 65535.f, e2 = (float)code[2] / 65535.f;
    const float* cp = d.params;
    uint32_t out[3];
    if (d.lut3n > 1) {
        // app 3D LUT: trilinear between the 8 nodes around the source colour
        const int n = d.lut3n;
        const float fr = fminf(fmaxf(e0, 0.0f), 1.0f) * (float)(n - 1);
        const float fg = fminf(fmaxf(e1, 0.0f), 1.0f) * (float)(n - 1);
        const float fb = fminf(fmaxf(e2, 0.0f), 1.0f) * (float)(n - 1);
        const int i0 = min((int)fr, n - 2), j0 = min((int)fg, n - 2), k0 = min((int)fb, n - 2);
        const float ar = fr - (float)i0, ag = fg - (float)j0, ab = fb - (float)k0;
        const float* L3 = d.lut3;
        auto at = [&](int ii, int jj, int kk, int c) { return L3[(((size_t)kk * n + jj) * n + ii) * 3 + c]; };
        for (int c = 0; c < 3; ++c) {
            const float c00 = at(i0, j0, k0, c) * (1.0f - ar) + at(i0 + 1, j0, k0, c) * ar;
            const float c10 = at(i0, j0 + 1, k0, c) * (1.0f - ar) + at(i0 + 1, j0 + 1, k0, c) * ar;
            const float c01 = at(i0, j0, k0 + 1, c) * (1.0f - ar) + at(i0 + 1, j0, k0 + 1, c) * ar;
            const float c11 = at(i0, j0 + 1, k0 + 1, c) * (1.0f - ar) + at(i0 + 1, j0 + 1, k0 + 1, c) * ar;
            const float v = (c00 * (1.0f - ag) + c10 * ag) * (1.0f - ab) + (c01 * (1.0f - ag) + c11 * ag) * ab;
            out[c] = (uint32_t)floorf(fminf(fmaxf(v, 0.0f), 1.0f) * 4095.0f + 0.5f);
        }
    } else {
        const float R = d.lut[lut_index(e0)], G = d.lut[lut_index(e1)], B = d.lut[lut_index(e2)];
        float X = cp[0] * R + cp[1] * G + cp[2] * B;
        float Y = cp[4] * R + cp[5] * G + cp[6] * B;
        float Z = cp[8] * R + cp[9] * G + cp[10] * B;
        if (cp[12] > 0.5f && Y > 0.0f) {  // tone map (PQ / HLG sources)
            float tr = cp[20] * X + cp[21] * Y + cp[22] * Z;
            float tg = cp[23] * X + cp[24] * Y + cp[25] * Z;
            float tb = cp[26] * X + cp[27] * Y + cp[28] * Z;
            const float pk = cp[17];
            if (cp[42] > 1.5f) {
                // each target-gamut channel; channels below 0 scale like the brightest one
                const float mx = fmaxf(tr, fmaxf(tg, tb)) * pk;
                const float sneg = mx > 0.0f ? tone_curve(cp, mx) / mx : 1.0f;
                tr = tr > 0.0f ? tone_curve(cp, tr * pk) / pk : tr * sneg;
                tg = tg > 0.0f ? tone_curve(cp, tg * pk) / pk : tg * sneg;
                tb = tb > 0.0f ? tone_curve(cp, tb * pk) / pk : tb * sneg;
                X = cp[29] * tr + cp[30] * tg + cp[31] * tb;
                Y = cp[32] * tr + cp[33] * tg + cp[34] * tb;
                Z = cp[35] * tr + cp[36] * tg + cp[37] * tb;
            } else {
                // luminance, or max(R,G,B) (hue kept); XYZ scaled alike
                const float m = cp[42] > 0.5f ? fmaxf(Y, fmaxf(tr, fmaxf(tg, tb)) * pk) : Y;
                const float s = tone_curve(cp, m) / m;
                X *= s;
                Y *= s;
                Z *= s;
            }
        }
        if (cp[18] > 0.5f) compress_gamut(cp, X, Y, Z);
        out[0] = encode_code(thr, X);
        out[1] = encode_code(thr, Y);
        out[2] = encode_code(thr, Z);
    }
    d.codes[i] = (uint16_t)out[0];
    d.codes[npix + i] = (uint16_t)out[1];
    d.codes[2 * (size_t)npix + i] = (uint16_t)out[2];
}

// ---------------------------------------------------------------- launchers
template <bool L0, int TX, int TY> static hipError_t launch_dwt(const DwtLevelDesc* descs, int n, int tiles, hipStream_t s) {
    hipLaunchKernelGGL((dwt_level_kernel<L0, TX, TY>), dim3(tiles), dim3(256), 0, s, descs, n);
    return hipGetLastError();
}

template <bool L0, int TY> static hipError_t launch_strip(const DwtLevelDesc* descs, int n, int tiles, hipStream_t s) {
    hipLaunchKernelGGL((dwt_strip_kernel<L0, TY>), dim3((tiles + 3) / 4), dim3(256), 0, s, descs, n, tiles);
    return hipGetLastError();
}

template <int TY, int PRE, int MINW>
static hipError_t launch_strip0(const DwtLevelDesc* descs, int n, int tiles, hipStream_t s) {
    hipLaunchKernelGGL((dwt_strip0_kernel<TY, PRE, MINW>), dim3((tiles + 3) / 4), dim3(256), 0, s, descs, n, tiles);
    return hipGetLastError();
}

template <int NW, int MINW>
static hipError_t launch_stage0(const DwtLevelDesc* descs, int n, int groups, hipStream_t s) {
    hipLaunchKernelGGL((dwt_stage0_kernel<NW, MINW>), dim3(groups), dim3(64 * NW), 0, s, descs, n, groups);
    return hipGetLastError();
}

template <int TY, int PRE, int MINW, int WG>
static hipError_t launch_strip0b(const DwtLevelDesc* descs, int n, int tiles, hipStream_t s) {
    hipLaunchKernelGGL((dwt_strip0b_kernel<TY, PRE, MINW, WG>), dim3((tiles + WG - 1) / WG), dim3(64 * WG), 0, s, descs, n, tiles);
    return hipGetLastError();
}

hipError_t launch_dwt_level(bool level0, int tile_shape, const DwtLevelDesc* descs, int ndesc, int tiles,
                            hipStream_t s) {
    switch (tile_shape) {
    case 0: return level0 ? launch_dwt<true, 64, 64>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 64>(descs, ndesc, tiles, s);
    case 1: return level0 ? launch_dwt<true, 64, 32>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 32>(descs, ndesc, tiles, s);
    case 2: return level0 ? launch_dwt<true, 32, 64>(descs, ndesc, tiles, s) : launch_dwt<false, 32, 64>(descs, ndesc, tiles, s);
    case 4: return level0 ? launch_strip<true, 64>(descs, ndesc, tiles, s) : launch_strip<false, 64>(descs, ndesc, tiles, s);
    case 5: return level0 ? launch_strip<true, 32>(descs, ndesc, tiles, s) : launch_strip<false, 32>(descs, ndesc, tiles, s);
    case 6: return level0 ? launch_strip<true, 128>(descs, ndesc, tiles, s) : launch_strip<false, 128>(descs, ndesc, tiles, s);
    case 7: return level0 ? launch_strip0<16, 10, 5>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 16>(descs, ndesc, tiles, s);
    case 8: return level0 ? launch_strip0<8, 15, 5>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 16>(descs, ndesc, tiles, s);
    case 9: return level0 ? launch_strip0b<8, 15, 5, 4>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 16>(descs, ndesc, tiles, s);
    case 10: return level0 ? launch_stage0<4, 5>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 16>(descs, ndesc, tiles, s);
    default: return level0 ? launch_dwt<true, 64, 16>(descs, ndesc, tiles, s) : launch_dwt<false, 64, 16>(descs, ndesc, tiles, s);
    }
}

void dwt_tile_size(int tile_shape, int& tx, int& ty) {
    static const int shapes[kDwtShapes][2] = {{64, 64}, {64, 32}, {32, 64}, {64, 16},
                                              {kStripOut, 64}, {kStripOut, 32}, {kStripOut, 128}, {kStripOut, 16},
                                              {kStripOut, 8}, {kStripOut, 8}, {4 * kStripOut, 8}};
    const int k = tile_shape >= 0 && tile_shape < kDwtShapes ? tile_shape : kDwtDefaultShape;
    tx = shapes[k][0];
    ty = shapes[k][1];
}

hipError_t launch_color(const ColorDesc& d, hipStream_t s) {
    const uint32_t npix = (uint32_t)d.width * d.height;
    hipLaunchKernelGGL(color_kernel, dim3((npix + 255) / 256), dim3(256), 0, s, d);
    return hipGetLastError();
}

}  // namespace gail_amd
