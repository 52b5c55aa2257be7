// PCRD-opt rate control on the GPU (rate_control.cpp:11-279 with the Tier-2 sizes of
// t2.cpp:31-254), between MQ coding and compaction, so the rate-control state stays in HBM
// and only the kept bytes of each code-block leave the GPU. The launches, in order:
//
//   rc_hull_kernel   one thread per code-block of every frame: the block's hull
//                    (rc_block_hull_to), stored point-major ([point][block]) so that the
//                    lanes of a wave read neighbouring blocks' points together, each point's
//                    slope bin floor(1024 log2 s), and per workgroup the smallest and largest
//                    finite slope;
//   rc_range_kernel  one workgroup per frame: the frame's slope range and histogram bins;
//   rc_fill_kernel   one thread per code-block: the slope histogram (bytes and included
//                    blocks per component, 32-bit integer atomics);
//   rc_scan_kernel   two passes over 1,024-bin segments: suffix sums of the histogram;
//   rc_kernel        one 1024-thread workgroup per frame: the lambda search of rc_shared.h
//                    in thread 0; each size measurement selects every block's passes (and
//                    writes them, with their length, over the block's pass count and
//                    codeword length, which compaction uses) and sizes every packet
//                    (rc_packet_bytes, in registers), the tile-part sums in LDS. At the end
//                    it clears the histogram bins it used, so every run finds them zero.
//
// The host library runs the same shared code, so both choose the same passes.
#include <hip/hip_runtime.h>

#include "device_types.h"
#include "kernels.h"
#include "rc_shared.h"

namespace gail_amd {

namespace {

constexpr int kHullThreads = 256;  // blocks per hull / fill workgroup (one partial slope range each)
constexpr int kScanThreads = 256;
constexpr int kScanBins = 4 * kScanThreads;  // bins per scan workgroup
constexpr int kScanSegs = (kRcMaxBins + 1 + kScanBins - 1) / kScanBins;
constexpr int kRcThreads = 1024;
constexpr int32_t kBinAbove = 0x7FFFFFFF;      // slope bin of an infinite slope
constexpr int32_t kBinBelow = -0x7FFFFFFF - 1;  // of a slope <= 0

__device__ __forceinline__ uint32_t* hist_of(const RcParams& P, uint32_t f, int a) {
    return P.hist + ((size_t)f * 6 + (size_t)a) * (kRcMaxBins + 1);
}

// Partial slope range of hull workgroup x of frame f: frames never share an entry.
__device__ __forceinline__ uint32_t part_index(const RcFrame& F, uint32_t f, uint32_t x) {
    return F.first_block / kHullThreads + f + x;
}

// RcSearch reads the histogram sums through data(c, b) and incl(c, b).
struct DevHistView {
    const uint32_t* rate[3];
    const uint32_t* incl[3];
    __device__ double data(int c, int b) const { return (double)rate[c][b]; }
    __device__ double incl_(int c, int b) const { return (double)incl[c][b]; }
};
struct DevHistRef {
    const DevHistView* h;
    __device__ double data(int c, int b) const { return h->data(c, b); }
    __device__ double incl(int c, int b) const { return h->incl_(c, b); }
};

// Frame-local view of the blocks for rc_packet_bytes: sel = passes << 24 | length.
struct DevBlocks {
    const uint32_t* sel;   // + first block
    const T1Info* info;    // + first block
    __device__ uint32_t passes(uint32_t i) const { return sel[i] >> 24; }
    __device__ uint32_t len(uint32_t i) const { return sel[i] & 0xFFFFFFu; }
    __device__ uint32_t nbps(uint32_t i) const { return info[i].nbps; }
};

// Point-major hull store of one block.
struct HullSink {
    double* slope;
    double* gain;
    uint32_t* sel;
    size_t stride;
    __device__ void operator()(int j, const RcPoint& q) const {
        const size_t at = (size_t)j * stride;
        slope[at] = q.slope;
        gain[at] = q.gain;
        sel[at] = q.passes << 24 | q.rate;
    }
};

}  // namespace

__global__ void __launch_bounds__(kHullThreads) rc_hull_kernel(RcParams P) {
    __shared__ unsigned long long s_min, s_max;
    const uint32_t f = blockIdx.y;
    const RcFrame& F = P.frames[f];
    if (F.max_total == 0) return;          // no rate control for this frame
    if (P.counters->flags & 3u) return;    // Tier-1 overflowed: the job is re-run
    if (blockIdx.x * kHullThreads >= F.nblocks) return;
    const int t = threadIdx.x;
    if (t == 0) {
        s_min = ~0ull;
        s_max = 0ull;
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * kHullThreads + t;
    unsigned long long lmin = ~0ull, lmax = 0ull;
    if (i < F.nblocks) {
        const uint32_t cb = F.first_block + i;
        const T1Info inf = P.info[cb];
        const int np = (int)inf.passes;
        const uint32_t band = P.blocks[cb].band;
        const int cnt = np ? rc_block_hull_to(np, (int)inf.nbps, P.band_scale[band], P.pass_len + (size_t)cb * kMaxPasses,
                                              P.pass_dist + (size_t)cb * kMaxPasses,
                                              HullSink{P.h_slope + cb, P.h_gain + cb, P.h_sel + cb, P.stride})
                           : 0;
        P.hull_n[cb] = (uint32_t)cnt | (P.bands[band].comp_mb & 3u) << 8;
        for (int h = 0; h < cnt; ++h) {
            const size_t at = (size_t)h * P.stride + cb;
            const double s = P.h_slope[at];
            int32_t k;
            if (!rc_finite(s)) k = kBinAbove;
            else if (!(s > 0)) k = kBinBelow;
            else {
                k = (int32_t)rc_slope_bin(s, P.mantissa, P.first);
                const unsigned long long bits = __builtin_bit_cast(unsigned long long, s);  // positive doubles order as their bits
                lmin = bits < lmin ? bits : lmin;
                lmax = bits > lmax ? bits : lmax;
            }
            P.h_kbin[at] = k;
        }
    }
    if (lmax) {
        atomicMin(&s_min, lmin);
        atomicMax(&s_max, lmax);
    }
    __syncthreads();
    if (t == 0) {
        unsigned long long* part = P.part + 2 * (size_t)part_index(F, f, blockIdx.x);
        part[0] = s_min;
        part[1] = s_max;
    }
}

__global__ void __launch_bounds__(kHullThreads) rc_range_kernel(RcParams P) {
    __shared__ unsigned long long s_min, s_max;
    const uint32_t f = blockIdx.x;
    RcFrame& F = P.frames[f];
    if (F.max_total == 0) return;
    if (P.counters->flags & 3u) return;
    const int t = threadIdx.x;
    if (t == 0) {
        s_min = ~0ull;
        s_max = 0ull;
    }
    __syncthreads();
    const uint32_t nparts = (F.nblocks + kHullThreads - 1) / kHullThreads;
    for (uint32_t x = t; x < nparts; x += kHullThreads) {
        const unsigned long long* part = P.part + 2 * (size_t)part_index(F, f, x);
        if (part[1]) {
            atomicMin(&s_min, part[0]);
            atomicMax(&s_max, part[1]);
        }
    }
    __syncthreads();
    if (t == 0) {
        double mn = __builtin_bit_cast(double, s_min), mx = __builtin_bit_cast(double, s_max);
        if (s_max == 0) mn = mx = 1.0;
        int64_t kLo;
        int nb;
        rc_bin_range(mn, mx, P.mantissa, P.first, kLo, nb);
        F.k_lo = kLo;
        F.nb = nb;
        F.min_slope = mn;
        F.max_slope = mx;
    }
}

__global__ void __launch_bounds__(kHullThreads) rc_fill_kernel(RcParams P) {
    const uint32_t f = blockIdx.y;
    const RcFrame& F = P.frames[f];
    if (F.max_total == 0) return;
    if (P.counters->flags & 3u) return;
    const uint32_t i = blockIdx.x * kHullThreads + threadIdx.x;
    if (i >= F.nblocks) return;
    const uint32_t cb = F.first_block + i;
    const int64_t kLo = F.k_lo;
    const int nb = F.nb;
    const uint32_t hn = P.hull_n[cb];
    const uint32_t cnt = hn & 0xFFu, c = hn >> 8;
    uint32_t* rate = hist_of(P, f, (int)c);
    uint32_t* incl = hist_of(P, f, 3 + (int)c);
    uint32_t prev = 0;
    for (uint32_t h = 0; h < cnt; ++h) {
        const size_t at = (size_t)h * P.stride + cb;
        const int64_t r = (int64_t)P.h_kbin[at] - kLo;
        const int b = (int)(r < 0 ? 0 : (r > nb - 1 ? nb - 1 : r));  // rc_bin_of on the stored bin
        const uint32_t rt = P.h_sel[at] & 0xFFFFFFu;
        atomicAdd(rate + b, rt - prev);
        if (h == 0) atomicAdd(incl + b, 1u);
        prev = rt;
    }
}

// Suffix sums of the 6 histogram arrays of each frame (v[b] += v[b + 1] + ... + v[nb]) in
// two passes over segments of kScanBins bins: PASS 0 stores each segment's total, PASS 1
// adds the totals of the segments above and scans its own segment in LDS.
template <int PASS>
__global__ void __launch_bounds__(kScanThreads) rc_scan_kernel(RcParams P) {
    __shared__ uint32_t s_v[kScanThreads];
    const uint32_t f = blockIdx.z;
    const int a = (int)blockIdx.y, seg = (int)blockIdx.x;
    const RcFrame& F = P.frames[f];
    if (F.max_total == 0) return;
    if (P.counters->flags & 3u) return;
    const int n1 = F.nb + 1;
    if (seg * kScanBins >= n1) return;
    const int t = threadIdx.x;
    uint32_t* v = hist_of(P, f, a);
    uint32_t* segsum = P.segsum + ((size_t)f * 6 + (size_t)a) * kScanSegs;
    const int b0 = seg * kScanBins + 4 * t;
    uint32_t x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = b0 + j < n1 ? v[b0 + j] : 0u;
    if (PASS == 0) {
        s_v[t] = x[0] + x[1] + x[2] + x[3];
        __syncthreads();
        for (int w = kScanThreads / 2; w > 0; w >>= 1) {
            if (t < w) s_v[t] += s_v[t + w];
            __syncthreads();
        }
        if (t == 0) segsum[seg] = s_v[0];
        return;
    }
    // totals of the segments above this one
    const int nseg = (n1 + kScanBins - 1) / kScanBins;
    uint32_t above = 0;
    if (t < 64)
        for (int s = seg + 1 + t; s < nseg; s += 64) above += segsum[s];
    // suffix sums inside the thread's four bins, then across the threads (Hillis-Steele)
    x[2] += x[3];
    x[1] += x[2];
    x[0] += x[1];
    s_v[t] = x[0];
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const uint32_t add = t + d < kScanThreads ? s_v[t + d] : 0u;
        __syncthreads();
        s_v[t] += add;
        __syncthreads();
    }
    __shared__ uint32_t s_above[64];
    if (t < 64) s_above[t] = above;
    __syncthreads();
    if (t == 0) {
        uint32_t sum = 0;
        for (int k = 0; k < 64; ++k) sum += s_above[k];
        s_above[0] = sum;
    }
    __syncthreads();
    const uint32_t carry = s_above[0] + (t + 1 < kScanThreads ? s_v[t + 1] : 0u);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (b0 + j < n1) v[b0 + j] = x[j] + carry;
}

__global__ void __launch_bounds__(kRcThreads) rc_kernel(RcParams P) {
    __shared__ unsigned long long s_tp[6];
    __shared__ double s_lambda[3];
    __shared__ double s_dist[kRcThreads];
    __shared__ int s_more;
    const uint32_t f = blockIdx.x;
    RcFrame& F = P.frames[f];
    if (F.max_total == 0) return;
    if (P.counters->flags & 3u) return;
    const int t = threadIdx.x;
    const uint32_t n = F.nblocks, b0 = F.first_block;
    const int nb = F.nb;

    RcSearch S;  // thread 0's
    if (t == 0) {
        S.init();
        S.kLo = F.k_lo;
        S.nb = nb;
        S.minSlope = F.min_slope;
        S.maxSlope = F.max_slope;
        S.maxTotal = F.max_total;
        S.maxComp = F.max_comp;
        S.mantissa = P.mantissa;
        S.first = P.first;
        for (int c = 0; c < 3; ++c) s_lambda[c] = 0.0;
    }
    __syncthreads();
    DevHistView H;
    for (int c = 0; c < 3; ++c) {
        H.rate[c] = hist_of(P, f, c);
        H.incl[c] = hist_of(P, f, 3 + c);
    }
    const DevHistRef HV{&H};
    const DevBlocks blk{P.sel + b0, P.info + b0};

    // ---- measure, search, measure ... (thread 0 decides, everyone measures)
    double dist = 0;
    for (;;) {
        // A hull's slopes fall strictly, so its points of slope >= lambda are a prefix and
        // the last of them (rc_select) is found without a data-dependent exit. The choice
        // and its length replace the block's pass count and codeword length at once (the
        // last measurement's stay); dist sums the gains of the points left out.
        dist = 0;
        for (uint32_t i = t; i < n; i += kRcThreads) {
            const uint32_t cb = b0 + i;
            const uint32_t hn = P.hull_n[cb];
            const uint32_t cnt = hn & 0xFFu;
            const double lam = s_lambda[hn >> 8];
            uint32_t sel = 0;
            for (uint32_t h = 0; h < cnt; h += 4) {  // four points' loads in flight at once
                double s[4], g[4];
                uint32_t hs[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const size_t at = (size_t)(h + j) * P.stride + cb;
                    const bool in = h + j < cnt;
                    s[j] = in ? P.h_slope[at] : 0.0;
                    g[j] = in ? P.h_gain[at] : 0.0;
                    hs[j] = in ? P.h_sel[at] : 0u;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (h + j < cnt) {
                        if (s[j] >= lam) sel = hs[j];
                        else dist += g[j];
                    }
            }
            P.sel[cb] = sel;
            P.info[cb].passes = sel >> 24;
            P.code_len[cb] = sel & 0xFFFFFFu;
        }
        if (t < 6) s_tp[t] = t < (int)F.ntp ? 14ull : 0ull;  // SOT + SOD of each tile-part
        __syncthreads();
        for (uint32_t k = t; k < F.npackets; k += kRcThreads) {
            const RcPacket pk = F.packets[k];
            atomicAdd(&s_tp[pk.tp], rc_packet_bytes(pk, F.pbands, blk));
        }
        __syncthreads();
        if (t == 0) {
            RcSizes sz;
            sz.total = F.fixed_bytes;
            sz.comp[0] = sz.comp[1] = sz.comp[2] = 0;
            for (uint32_t k = 0; k < F.ntp; ++k) {
                sz.comp[F.tp_comp[k]] += s_tp[k];
                sz.total += s_tp[k];
            }
            const bool more = S.next(HV, sz);
            s_more = more ? 1 : 0;
            if (more)
                for (int c = 0; c < 3; ++c) s_lambda[c] = S.lambda[c];
            else {
                for (int c = 0; c < 3; ++c) F.lambda[c] = S.lambda[c];
                F.total = sz.total;
                for (int c = 0; c < 3; ++c) F.comp[c] = sz.comp[c];
                for (uint32_t k = 0; k < 6; ++k) F.tile_part[k] = k < F.ntp ? s_tp[k] : 0ull;
                F.iterations = S.iterations;
                F.status = 0;
            }
        }
        __syncthreads();
        if (!s_more) break;
    }

    // ---- distortion left (fixed-order tree sum), and a clean histogram for the next run
    s_dist[t] = dist;
    __syncthreads();
    for (int w = kRcThreads / 2; w > 0; w >>= 1) {
        if (t < w) s_dist[t] += s_dist[t + w];
        __syncthreads();
    }
    if (t == 0) F.distortion = s_dist[0];
    for (int a = 0; a < 6; ++a) {
        uint32_t* v = hist_of(P, f, a);
        for (int b = t; b <= nb; b += kRcThreads) v[b] = 0u;
    }
}

hipError_t launch_rc(const RcParams& p, int frames, uint32_t max_frame_blocks, hipStream_t s) {
    const uint32_t gx = (max_frame_blocks + kHullThreads - 1) / kHullThreads;
    if (!gx || !frames) return hipSuccess;
    hipLaunchKernelGGL(rc_hull_kernel, dim3(gx, frames), dim3(kHullThreads), 0, s, p);
    hipLaunchKernelGGL(rc_range_kernel, dim3(frames), dim3(kHullThreads), 0, s, p);
    hipLaunchKernelGGL(rc_fill_kernel, dim3(gx, frames), dim3(kHullThreads), 0, s, p);
    hipLaunchKernelGGL(rc_scan_kernel<0>, dim3(kScanSegs, 6, frames), dim3(kScanThreads), 0, s, p);
    hipLaunchKernelGGL(rc_scan_kernel<1>, dim3(kScanSegs, 6, frames), dim3(kScanThreads), 0, s, p);
    hipLaunchKernelGGL(rc_kernel, dim3(frames), dim3(kRcThreads), 0, s, p);
    return hipGetLastError();
}

size_t rc_segsum_entries(size_t frames) { return frames * 6 * kScanSegs; }

}  // namespace gail_amd
