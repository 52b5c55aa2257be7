// PCRD-opt rate control on the GPU (rate_control.cpp:11-279 with the Tier-2 sizes of
// t2.cpp:31-254), one workgroup per frame of a job, between MQ coding and compaction, so the
// rate-control state stays in HBM and only the kept bytes of each code-block leave the GPU.
//
// The workgroup builds every block's hull (rc_block_hull), the frame's slope histogram on
// the exact 1/1024-octave grid (integer atomics, then suffix sums), and runs the lambda
// search of rc_shared.h in thread 0; each size measurement selects every block's passes and
// sizes every packet in parallel (rc_packet_bytes), the tile-part sums in LDS. The chosen
// passes then overwrite each block's pass count and codeword length, which compaction uses.
// The host library runs the same shared code, so both choose the same passes.
#include <hip/hip_runtime.h>

#include "device_types.h"
#include "kernels.h"
#include "rc_shared.h"

namespace gail_amd {

namespace {

constexpr int kRcThreads = 256;

struct DevHist {
    const unsigned long long* rate[3];
    const unsigned long long* incl[3];
    __device__ double data(int c, int b) const { return (double)rate[c][b]; }
    __device__ double incl_(int c, int b) const { return (double)incl[c][b]; }
};

// RcSearch reads the histogram through data(c, b) and incl(c, b).
struct DevHistView {
    const DevHist* h;
    __device__ double data(int c, int b) const { return h->data(c, b); }
    __device__ double incl(int c, int b) const { return h->incl_(c, b); }
};

// Frame-local view of the blocks for rc_packet_bytes.
struct DevBlocks {
    const uint8_t* chosen;     // + first block
    const T1Info* info;        // + first block
    const uint32_t* pass_len;  // + first block * kMaxPasses
    __device__ uint32_t passes(uint32_t i) const { return chosen[i]; }
    __device__ uint32_t nbps(uint32_t i) const { return info[i].nbps; }
    __device__ uint32_t len(uint32_t i) const {
        const uint32_t np = chosen[i];
        return np ? pass_len[(size_t)i * kMaxPasses + np - 1] : 0u;
    }
};

__device__ __forceinline__ void atomic_min_u64(unsigned long long* a, unsigned long long v) { atomicMin(a, v); }
__device__ __forceinline__ void atomic_max_u64(unsigned long long* a, unsigned long long v) { atomicMax(a, v); }

// In-place suffix sums of v[0..n] (v[b] += v[b + 1] + ... + v[n]) by the whole workgroup.
__device__ void suffix_sums(unsigned long long* v, int n1, unsigned long long* s_part) {
    const int t = threadIdx.x;
    const int chunk = (n1 + kRcThreads - 1) / kRcThreads;
    const int b0 = t * chunk, b1 = b0 + chunk < n1 ? b0 + chunk : n1;
    unsigned long long acc = 0;
    for (int b = b1 - 1; b >= b0; --b) {
        acc += v[b];
        v[b] = acc;
    }
    s_part[t] = b0 < b1 ? acc : 0ull;
    __syncthreads();
    if (t == 0) {  // exclusive suffix sums of the chunk totals
        unsigned long long run = 0;
        for (int k = kRcThreads - 1; k >= 0; --k) {
            const unsigned long long x = s_part[k];
            s_part[k] = run;
            run += x;
        }
    }
    __syncthreads();
    const unsigned long long add = s_part[t];
    for (int b = b0; b < b1; ++b) v[b] += add;
    __syncthreads();
}

}  // namespace

__global__ void __launch_bounds__(kRcThreads) rc_kernel(RcParams P) {
    __shared__ unsigned long long s_min, s_max;
    __shared__ unsigned long long s_tp[6];
    __shared__ unsigned long long s_part[kRcThreads];
    __shared__ double s_lambda[3];
    __shared__ double s_dist[kRcThreads];
    __shared__ int64_t s_klo;
    __shared__ int s_nb, s_more;
    RcFrame& F = P.frames[blockIdx.x];
    if (F.max_total == 0) return;          // no rate control for this frame
    if (P.counters->flags & 3u) return;    // Tier-1 overflowed: the job is re-run
    const int t = threadIdx.x;
    const uint32_t n = F.nblocks, b0 = F.first_block;
    unsigned long long* hist = P.hist + (size_t)blockIdx.x * 6 * (kRcMaxBins + 1);

    // ---- hulls, and the frame's smallest and largest finite slopes
    if (t == 0) {
        s_min = ~0ull;
        s_max = 0ull;
    }
    __syncthreads();
    unsigned long long lmin = ~0ull, lmax = 0ull;
    for (uint32_t i = t; i < n; i += kRcThreads) {
        const uint32_t cb = b0 + i;
        const T1Info inf = P.info[cb];
        const int np = (int)inf.passes;
        RcPoint* pts = P.hull + (size_t)cb * kMaxPasses;
        const int cnt = np ? rc_block_hull(np, (int)inf.nbps, P.band_scale[P.blocks[cb].band],
                                           P.pass_len + (size_t)cb * kMaxPasses, P.pass_dist + (size_t)cb * kMaxPasses, pts)
                           : 0;
        P.hull_n[cb] = (uint32_t)cnt;
        for (int h = 0; h < cnt; ++h) {
            const double s = pts[h].slope;
            if (rc_finite(s) && s > 0) {  // positive doubles order as their bits
                const unsigned long long bits = __builtin_bit_cast(unsigned long long, s);
                lmin = bits < lmin ? bits : lmin;
                lmax = bits > lmax ? bits : lmax;
            }
        }
    }
    if (lmax) {
        atomic_min_u64(&s_min, lmin);
        atomic_max_u64(&s_max, lmax);
    }
    __syncthreads();
    RcSearch S;  // thread 0's
    if (t == 0) {
        double mn = __builtin_bit_cast(double, s_min), mx = __builtin_bit_cast(double, s_max);
        if (s_max == 0) mn = mx = 1.0;
        S.init();
        rc_bin_range(mn, mx, P.mantissa, P.first, S.kLo, S.nb);
        S.minSlope = mn;
        S.maxSlope = mx;
        S.maxTotal = F.max_total;
        S.maxComp = F.max_comp;
        S.mantissa = P.mantissa;
        S.first = P.first;
        s_klo = S.kLo;
        s_nb = S.nb;
        F.min_slope = mn;
        F.max_slope = mx;
        for (int c = 0; c < 3; ++c) s_lambda[c] = 0.0;
    }
    __syncthreads();
    const int64_t kLo = s_klo;
    const int nb = s_nb;

    // ---- slope histogram: bytes and included blocks per component, then suffix sums
    const int n1 = nb + 1;
    for (int k = t; k < 6 * n1; k += kRcThreads) hist[(size_t)(k / n1) * (kRcMaxBins + 1) + k % n1] = 0ull;
    __syncthreads();
    for (uint32_t i = t; i < n; i += kRcThreads) {
        const uint32_t cb = b0 + i;
        const int c = (int)(P.bands[P.blocks[cb].band].comp_mb & 3u);
        unsigned long long* rate = hist + (size_t)c * (kRcMaxBins + 1);
        unsigned long long* incl = hist + (size_t)(3 + c) * (kRcMaxBins + 1);
        const RcPoint* pts = P.hull + (size_t)cb * kMaxPasses;
        const uint32_t cnt = P.hull_n[cb];
        uint32_t prev = 0;
        for (uint32_t h = 0; h < cnt; ++h) {
            const int b = rc_bin_of(pts[h].slope, kLo, nb, P.mantissa, P.first);
            atomicAdd(rate + b, (unsigned long long)(pts[h].rate - prev));
            if (h == 0) atomicAdd(incl + b, 1ull);
            prev = pts[h].rate;
        }
    }
    __syncthreads();
    for (int a = 0; a < 6; ++a) suffix_sums(hist + (size_t)a * (kRcMaxBins + 1), n1, s_part);
    DevHist H;
    for (int c = 0; c < 3; ++c) {
        H.rate[c] = hist + (size_t)c * (kRcMaxBins + 1);
        H.incl[c] = hist + (size_t)(3 + c) * (kRcMaxBins + 1);
    }
    const DevHistView HV{&H};
    const DevBlocks blk{P.chosen + b0, P.info + b0, P.pass_len + (size_t)b0 * kMaxPasses};

    // ---- measure, search, measure ... (thread 0 decides, everyone measures)
    RcTagTree incl, zbp;
    for (;;) {
        for (uint32_t i = t; i < n; i += kRcThreads) {
            const uint32_t cb = b0 + i;
            const int c = (int)(P.bands[P.blocks[cb].band].comp_mb & 3u);
            P.chosen[cb] = (uint8_t)rc_select(P.hull + (size_t)cb * kMaxPasses, (int)P.hull_n[cb], s_lambda[c]);
        }
        if (t < 6) s_tp[t] = t < (int)F.ntp ? 14ull : 0ull;  // SOT + SOD of each tile-part
        __syncthreads();
        for (uint32_t k = t; k < F.npackets; k += kRcThreads) {
            const RcPacket pk = F.packets[k];
            atomicAdd(&s_tp[pk.tp], rc_packet_bytes(pk, F.pbands, blk, incl, zbp));
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

    // ---- keep the chosen passes: compaction copies only those, and only their bytes
    double dist = 0;
    for (uint32_t i = t; i < n; i += kRcThreads) {
        const uint32_t cb = b0 + i;
        const uint32_t np = P.chosen[cb];
        const RcPoint* pts = P.hull + (size_t)cb * kMaxPasses;
        for (uint32_t h = 0; h < P.hull_n[cb]; ++h)
            if (pts[h].passes > np) dist += pts[h].gain;
        P.info[cb].passes = np;
        P.code_len[cb] = np ? P.pass_len[(size_t)cb * kMaxPasses + np - 1] : 0u;
    }
    s_dist[t] = dist;
    __syncthreads();
    for (int w = kRcThreads / 2; w > 0; w >>= 1) {  // fixed-order tree sum
        if (t < w) s_dist[t] += s_dist[t + w];
        __syncthreads();
    }
    if (t == 0) F.distortion = s_dist[0];
}

hipError_t launch_rc(const RcParams& p, int frames, hipStream_t s) {
    hipLaunchKernelGGL(rc_kernel, dim3(frames), dim3(kRcThreads), 0, s, p);
    return hipGetLastError();
}

}  // namespace gail_amd
