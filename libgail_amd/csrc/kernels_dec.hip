// Decoder kernels for CDNA4 (gfx950, wave64).
//
// Tier-1: the frame's longest code-blocks get a wave each, whose lanes all run the same
// values, so the decode runs on scalar registers and branches; the rest go one lane per block
// (t1_decode_block.h), ordered longest first and packed into waves of at most 16 blocks by a
// byte budget on the host so a wave's lanes finish together. A lane's row masks and context states live in LDS at [entry][slot] (conflict-free
// whatever row each lane is on); only 16 lanes of each wave hold a slot, so four waves share
// the LDS of one and a SIMD keeps several waves to hide the serial chain's latencies. The
// codeword comes through a 16-byte register window, and the reconstruction words go out as
// atomic adds into zeroed Mallat planes, so no lane ever waits on a read of its own output.
//
// Inverse DWT: one launch per level; a workgroup owns a TX x TY tile of the level's output
// and all three components. It gathers the tile plus a 4-sample mirrored halo from the four
// quadrants (dequantising the high bands, and the top level's LL, from their words on load),
// lifts the rows, then the columns, in registers, and writes the interleaved tile with
// coalesced row stores. Every sample sees the CPU reference's operations in its order
// (rows, then columns; scaling, then lifting steps -delta, -gamma, -beta, -alpha), and the
// kernels are built with -ffp-contract=off, so results are bit-identical to it.
//
// Inverse ICT: one thread per pixel, level shift, rounding, clamping to 12 bits, packing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels_dec.h"
#include "t1_decode_block.h"
#include "t1_tables.h"

namespace gail_amd {

namespace {

__constant__ __attribute__((aligned(16))) uint8_t kDecZc[4 * 128];
__constant__ uint32_t kDecMq[47];

// ISO/IEC 15444-1 Table F.4, rounded to float as the CPU reference rounds them.
constexpr float kK = (float)1.230174104914001;
constexpr float kInvK = (float)(1.0 / 1.230174104914001);
constexpr float kNegAlpha = (float)1.586134342059924;
constexpr float kNegBeta = (float)0.052980118572961;
constexpr float kNegGamma = (float)-0.882911075530934;
constexpr float kNegDelta = (float)-0.443506852043971;
// Inverse ICT (G.3).
constexpr float kR2 = (float)1.402, kG1 = (float)-0.34413, kG2 = (float)-0.71414, kB1 = (float)1.772;

// ---------------------------------------------------------------- Tier-1
struct LdsState {
    uint32_t (*masks)[32][kDecSlots];
    uint32_t (*ctxs)[kDecSlots];
    const uint32_t* mq;   // Table C.2 in LDS
    const uint8_t* zct;   // Table D.1 of the block's orientation in LDS
    int lane;
    __device__ __forceinline__ void reset() {
        for (uint32_t r = 0; r < 32u; ++r)
#pragma unroll
            for (int t = 0; t < 4; ++t) masks[t][r][lane] = 0u;
        for (uint32_t c = 0; c < 19u; ++c) ctxs[c][lane] = mq[t1d_initial_state(c)];
    }
    __device__ __forceinline__ uint32_t rows(int t, uint32_t r) const { return masks[t][r][lane]; }
    __device__ __forceinline__ void set_rows(int t, uint32_t r, uint32_t v) { masks[t][r][lane] = v; }
    __device__ __forceinline__ uint32_t ctx(uint32_t c) const { return ctxs[c][lane]; }
    __device__ __forceinline__ void set_ctx(uint32_t c, uint32_t v) { ctxs[c][lane] = v; }
    __device__ __forceinline__ uint32_t tab(uint32_t i) const { return mq[i]; }
    __device__ __forceinline__ uint32_t zc(uint32_t code) const { return zct[code]; }
};

struct AtomicSink {
    uint32_t* base;
    uint32_t stride;
    __device__ __forceinline__ void add(uint32_t i, uint32_t v) {
        __hip_atomic_fetch_add(base + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// A whole wave on one block: every decoding value is the same in all lanes, so the compiler
// keeps it in scalar registers and the control flow on scalar branches (no exec-mask work).
// The block's state lives in vector registers used as per-lane cells, read and written with
// readlane / writelane (a few cycles, no LDS round trip on the symbol chain): lane r of m[t]
// holds row mask r of set t, lane k of c context k; lane i of mqv holds Table C.2 entry i,
// lane j of zcv the zero-coding contexts of codes 4j..4j+3 for the block's orientation. A write
// is a select on the lane number. Lane 0 alone adds to the words.
struct WaveRegState {
    uint32_t m0, m1, m2, m3, c, mqv, zcv, lid;
    __device__ __forceinline__ void reset() {
        m0 = m1 = m2 = m3 = 0u;
        c = kDecMq[t1d_initial_state(threadIdx.x & 63)];  // lanes past 18 hold unused states
    }
    __device__ __forceinline__ uint32_t& set(int t) { return t == 0 ? m0 : (t == 1 ? m1 : (t == 2 ? m2 : m3)); }
    __device__ __forceinline__ uint32_t rows(int t, uint32_t r) const {
        return (uint32_t)__builtin_amdgcn_readlane((int)(t == 0 ? m0 : (t == 1 ? m1 : (t == 2 ? m2 : m3))), (int)r);
    }
    __device__ __forceinline__ void set_rows(int t, uint32_t r, uint32_t v) {
        uint32_t& m = set(t);
        m = lid == r ? v : m;
    }
    __device__ __forceinline__ uint32_t ctx(uint32_t k) const { return (uint32_t)__builtin_amdgcn_readlane((int)c, (int)k); }
    __device__ __forceinline__ void set_ctx(uint32_t k, uint32_t v) { c = lid == k ? v : c; }
    __device__ __forceinline__ uint32_t tab(uint32_t i) const { return (uint32_t)__builtin_amdgcn_readlane((int)mqv, (int)i); }
    __device__ __forceinline__ uint32_t zc(uint32_t code) const {
        return ((uint32_t)__builtin_amdgcn_readlane((int)zcv, (int)(code >> 2)) >> (8u * (code & 3u))) & 0xFFu;
    }
};

struct WaveSink {
    uint32_t* base;
    uint32_t stride;
    __device__ __forceinline__ void add(uint32_t i, uint32_t v) {
        if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(base + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

__device__ __forceinline__ bool block_fields(const T1DecodeParams& P, const DecBlock& d, uint32_t& w, uint32_t& h,
                                             uint32_t& orient, uint32_t& nbps, uint32_t& passes) {
    w = (d.geom & 31u) + 1u;
    h = ((d.geom >> 5) & 31u) + 1u;
    orient = (d.geom >> 10) & 3u;
    nbps = (d.geom >> 12) & 63u;
    passes = (d.geom >> 18) & 255u;
    // the host builds the table from the layout; a wrong one still writes nothing outside
    return d.dst < P.nwords && P.nwords - d.dst >= (h - 1u) * P.stride + w;
}

__global__ void __launch_bounds__(64 * kDecWaves) t1_decode_kernel(T1DecodeParams P) {
    __shared__ uint32_t s_rows[4][32][kDecSlots];
    __shared__ uint32_t s_ctx[19][kDecSlots];
    __shared__ uint32_t s_tab[48];
    __shared__ uint8_t s_zc[4 * 128];
    const int t = threadIdx.x;
    if (t < 47) s_tab[t] = kDecMq[t];
    for (int i = t; i < 4 * 128; i += 64 * kDecWaves) s_zc[i] = kDecZc[i];
    __syncthreads();
    const int lane = t & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t gw = blockIdx.x * kDecWaves + wave;
    uint32_t w, h, orient, nbps, passes;
    if (gw < P.nlong) {
        // one of the frame's longest blocks: the whole wave decodes it
        const DecBlock d = P.blocks[gw];
        if (d.dst == kDecEmpty || !block_fields(P, d, w, h, orient, nbps, passes)) return;
        WaveRegState st;
        st.lid = (uint32_t)lane;
        st.mqv = lane < 47 ? kDecMq[lane] : 0u;
        st.zcv = lane < 32 ? reinterpret_cast<const uint32_t*>(kDecZc + 128u * orient)[lane] : 0u;
        WaveSink out{P.words + d.dst, P.stride};
        t1d_decode_block(st, P.bytes, d.off, d.len, w, h, nbps, passes, out);
        return;
    }
    if (lane >= kDecWaveLanes) return;
    const uint32_t i = P.nlong + (gw - P.nlong) * kDecWaveLanes + lane;
    if (i >= P.nblocks) return;
    const DecBlock d = P.blocks[i];
    if (d.dst == kDecEmpty || !block_fields(P, d, w, h, orient, nbps, passes)) return;
    LdsState st{s_rows, s_ctx, s_tab, s_zc + 128u * orient, (int)wave * kDecWaveLanes + lane};
    AtomicSink out{P.words + d.dst, P.stride};
    t1d_decode_block(st, P.bytes, d.off, d.len, w, h, nbps, passes, out);
}

// ---------------------------------------------------------------- inverse DWT
// Whole-sample symmetric extension of index i into [0, n).
__device__ __forceinline__ int mirror(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (n == 1) return 0;
    const int period = 2 * (n - 1);
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

__device__ __forceinline__ float dequant(uint32_t wd, float half) {
    const float a = (float)(wd & 0x7FFFFFFFu) * half;
    return (wd & 0x80000000u) ? -a : a;
}

// Inverse lifting of N samples in registers, sample k at parity k & 1 (the tile origin is
// even). The end samples lack a neighbour and are left alone; the halo absorbs the damage.
template <int N> __device__ __forceinline__ void unlift97(float (&v)[N]) {
#pragma unroll
    for (int k = 2; k + 1 < N; k += 2) v[k] += kNegDelta * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 1; k + 1 < N; k += 2) v[k] += kNegGamma * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 2; k + 1 < N; k += 2) v[k] += kNegBeta * (v[k - 1] + v[k + 1]);
#pragma unroll
    for (int k = 1; k + 1 < N; k += 2) v[k] += kNegAlpha * (v[k - 1] + v[k + 1]);
}

constexpr int kIdwtTX = 64, kIdwtTY = 32;

template <int TX, int TY> __global__ void __launch_bounds__(256) idwt_level_kernel(IdwtLevel P) {
    constexpr int RX = TX + 8, RY = TY + 8, PITCH = RX + 1;
    constexpr int NPOS = RX * RY;
    constexpr int PER = (NPOS + 255) / 256;
    __shared__ float tile[3][RY][PITCH];
    // Workgroup b runs on XCD b % 8; each XCD takes a contiguous range of tiles, so tiles
    // that share halo rows meet in the same L2.
    const int nb = (int)gridDim.x, q = nb >> 3, rem = nb & 7;
    const int xcd = (int)blockIdx.x & 7, jb = (int)blockIdx.x >> 3;
    const int t = xcd * q + min(xcd, rem) + jb;
    const int x0 = (t % P.tiles_x) * TX, y0 = (t / P.tiles_x) * TY;
    const int w = P.w, h = P.h, sw = (w + 1) >> 1, sh = (h + 1) >> 1;
    const bool rows_on = w >= 2, cols_on = h >= 2;
    const int tid = threadIdx.x;

    // Gather: the interleaved sample (gx, gy) is column gx / 2 of the low or high half and
    // row gy / 2 of the low or high half, by parity.
    uint32_t at[PER];
    int orient[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = tid + 256 * j;
        const int ry = i / RX, rx = i - ry * RX;
        const int gy = mirror(y0 - 4 + ry, h), gx = mirror(x0 - 4 + rx, w);
        const int r = (gy & 1) ? sh + (gy >> 1) : (gy >> 1);
        const int c = (gx & 1) ? sw + (gx >> 1) : (gx >> 1);
        at[j] = (uint32_t)r * P.stride + (uint32_t)c;
        orient[j] = (gx & 1) | ((gy & 1) << 1);
    }
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
        const uint32_t* wsrc = P.words + (size_t)comp * P.plane;
        const float* lsrc = P.ll ? P.ll + (size_t)comp * P.plane : nullptr;
        const float h0 = P.half[comp][0], h1 = P.half[comp][1], h2 = P.half[comp][2], h3 = P.half[comp][3];
        float v[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (NPOS % 256 != 0 && tid + 256 * j >= NPOS) continue;
            const int o = orient[j];
            if (o == 0 && lsrc) {
                v[j] = lsrc[at[j]];
            } else {
                v[j] = dequant(wsrc[at[j]], o == 0 ? h0 : (o == 1 ? h1 : (o == 2 ? h2 : h3)));
            }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = tid + 256 * j;
            if (NPOS % 256 != 0 && i >= NPOS) continue;
            const int ry = i / RX, rx = i - ry * RX;
            // the row transform starts by scaling: low half times K, high half times 1/K
            tile[comp][ry][rx] = rows_on ? v[j] * ((rx & 1) ? kInvK : kK) : v[j];
        }
    }
    __syncthreads();
    // Rows: one thread per staged row of each component.
    if (rows_on) {
        for (int i = tid; i < 3 * RY; i += 256) {
            const int c = i / RY, ry = i - c * RY;
            float v[RX];
#pragma unroll
            for (int k = 0; k < RX; ++k) v[k] = tile[c][ry][k];
            unlift97(v);
#pragma unroll
            for (int k = 4; k < TX + 4; ++k) tile[c][ry][k] = v[k];
        }
        __syncthreads();
    }
    // Columns: one thread per output column of each component.
    if (cols_on) {
        for (int i = tid; i < 3 * TX; i += 256) {
            const int c = i / TX, rx = 4 + (i - c * TX);
            float v[RY];
#pragma unroll
            for (int k = 0; k < RY; ++k) v[k] = tile[c][k][rx] * ((k & 1) ? kInvK : kK);
            unlift97(v);
#pragma unroll
            for (int k = 4; k < TY + 4; ++k) tile[c][k][rx] = v[k];
        }
        __syncthreads();
    }
    // Coalesced row stores of the interleaved tile.
    constexpr int ROWS = 256 / TX;
    const int k = tid % TX, r0 = tid / TX, gx = x0 + k;
    if (gx >= w) return;
#pragma unroll 4
    for (int j = 0; j < 3 * TY / ROWS; ++j) {
        const int c = j / (TY / ROWS), r = r0 + ROWS * (j % (TY / ROWS)), gy = y0 + r;
        if (gy < h) P.dst[(size_t)c * P.plane + (size_t)gy * P.stride + gx] = tile[c][4 + r][4 + k];
    }
}

// Dequantised words of a whole frame (the wavelet planes before the inverse DWT).
__global__ void __launch_bounds__(256) dequant_frame_kernel(DequantFrame P) {
    __shared__ float s_half[9 * 4];
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, c = blockIdx.z;
    if (threadIdx.x < 36) s_half[threadIdx.x] = P.half[c][threadIdx.x >> 2][threadIdx.x & 3];
    __syncthreads();
    if (x >= P.width) return;
    // the band holding (x, y) in the Mallat layout
    uint32_t ww = P.width, hh = P.height, level = P.levels, orient = 0;
    for (uint32_t l = 1; l <= P.levels; ++l) {
        const uint32_t nw = (ww + 1) >> 1, nh = (hh + 1) >> 1;
        if (x >= nw || y >= nh) {
            level = l;
            orient = (x >= nw ? 1u : 0u) | (y >= nh ? 2u : 0u);
            break;
        }
        ww = nw;
        hh = nh;
    }
    const size_t i = (size_t)c * P.width * P.height + (size_t)y * P.width + x;
    P.dst[i] = dequant(P.words[i], s_half[level * 4 + orient]);
}

// ---------------------------------------------------------------- inverse ICT and packing
__global__ void __launch_bounds__(256) ict_pack_kernel(IctPack P) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= P.w) return;
    const size_t i = (size_t)y * P.stride + x;
    const float y0 = P.planes[i], y1 = P.planes[P.plane + i], y2 = P.planes[2 * (size_t)P.plane + i];
    float v[3];
    v[0] = y0 + kR2 * y2;
    v[1] = y0 + kG1 * y1 + kG2 * y2;
    v[2] = y0 + kB1 * y1;
    uint16_t* o = reinterpret_cast<uint16_t*>(P.out + (int64_t)y * P.pitch) + 3 * (size_t)x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int code = (int)floorf(v[k] + 2048.5f);
        code = code < 0 ? 0 : (code > 4095 ? 4095 : code);
        o[k] = (uint16_t)(code << P.shift);
    }
}

}  // namespace

hipError_t dec_tables_init() {
    uint8_t z[4 * 128];
    build_zc_table(z);
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kDecZc), z, sizeof z);
    if (e != hipSuccess) return e;
    uint32_t tab[47];
    build_mq_table(tab);
    return hipMemcpyToSymbol(HIP_SYMBOL(kDecMq), tab, sizeof tab);
}

// Zeroes n words at p (16-byte aligned): four words per store, a grid-stride loop.
__global__ void __launch_bounds__(256) zero_words_kernel(uint32_t* p, uint64_t n) {
    const uint64_t quads = n / 4, step = (uint64_t)gridDim.x * blockDim.x;
    uint4* q = reinterpret_cast<uint4*>(p);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < quads; i += step) q[i] = make_uint4(0, 0, 0, 0);
    const uint64_t t = quads * 4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && t < n) p[t] = 0u;
}

hipError_t launch_zero_words(uint32_t* p, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    if ((uintptr_t)p % 16) return hipErrorInvalidValue;
    const uint64_t quads = n / 4;
    const unsigned groups = (unsigned)std::min<uint64_t>(4096, (quads + 255) / 256 + 1);
    hipLaunchKernelGGL(zero_words_kernel, dim3(groups), dim3(256), 0, s, p, (uint64_t)n);
    return hipGetLastError();
}

hipError_t launch_t1_decode(const T1DecodeParams& p, hipStream_t s) {
    if (!p.nblocks) return hipSuccess;
    const uint32_t groups = (t1_decode_waves(p.nblocks, p.nlong) + kDecWaves - 1) / kDecWaves;
    hipLaunchKernelGGL(t1_decode_kernel, dim3(groups), dim3(64 * kDecWaves), 0, s, p);
    return hipGetLastError();
}

void idwt_tile_size(int& tx, int& ty) {
    tx = kIdwtTX;
    ty = kIdwtTY;
}

hipError_t launch_idwt_level(const IdwtLevel& p, hipStream_t s) {
    hipLaunchKernelGGL((idwt_level_kernel<kIdwtTX, kIdwtTY>), dim3(p.tiles), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_dequant_frame(const DequantFrame& p, hipStream_t s) {
    hipLaunchKernelGGL(dequant_frame_kernel, dim3((p.width + 255) / 256, p.height, 3), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_ict_pack(const IctPack& p, hipStream_t s) {
    hipLaunchKernelGGL(ict_pack_kernel, dim3((p.w + 255) / 256, p.h), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace gail_amd
