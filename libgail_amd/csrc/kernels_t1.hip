// EBCOT Tier-1 for CDNA4 (gfx950, wave64): fused dead-zone quantisation + context
// modelling, block ordering, MQ coding with exact truncation lengths, and compaction.
//
// Context modelling: one wave codes two 32x32 code-blocks, lanes 0-31 and 32-63; lane x
// owns column x and keeps its column's state as 32-bit row masks (significant, refined,
// visited, current magnitude bit) plus its 32 quantised magnitudes in registers. Left and
// right neighbours come from lane shuffles, above and below are bit shifts. Per 4-row
// stripe the significance-propagation pass is solved as a monotone fixpoint over the
// half-wave (no change anywhere => the sequential scan's result), cleanup needs no
// fixpoint because every remaining sample is coded. Each lane packs its column's
// symbols of a stripe (at most 10, 6 bits each) in a 64-bit register; ballots + mbcnt
// give the exclusive prefix over the columns, so symbols land in scan order in a per-block
// LDS staging buffer, copied to an HBM pool once per bit-plane.
//
// MQ coding: one lane per code-block, blocks ordered largest first so a wave's lanes
// finish together. The coder keeps A, C, CT and the pending byte in registers and the 19
// context states in LDS; at each pass end it snapshots its state, and after the final
// flush it derives every pass's truncation length from the snapshots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device_types.h"
#include "kernels.h"
#include "t1_cm_masks.h"
#include "t1_tables.h"

namespace gail_amd {

namespace {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of the lane one column to the left / right within the same half-wave; zero at
// the block's edges.
// DPP wave_shr:1 / wave_shl:1 move a register one lane up / down across the whole wave in
// one VALU operation; the lanes at a block's edge (x = 0 / 31) see zero.
__device__ __forceinline__ uint32_t from_left(uint32_t v, int x) {
    const uint32_t r = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);
    return x == 0 ? 0u : r;
}
__device__ __forceinline__ uint32_t from_right(uint32_t v, int x) {
    const uint32_t r = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);
    return x == 31 ? 0u : r;
}

__device__ __forceinline__ double half_sum_d(double v) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 32);
    return v;
}
__device__ __forceinline__ uint32_t half_max_u(uint32_t v) {
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        const uint32_t t = (uint32_t)__shfl_xor((int)v, o, 32);
        v = t > v ? t : v;
    }
    return v;
}

// Inclusive prefix sum over the lanes of each half-wave (lanes 0-31 and 32-63 separately):
// four row shifts within each 16-lane row, then row 0's (row 2's) last lane added to row 1
// (row 3). Lanes without a source add zero.
__device__ __forceinline__ uint32_t half_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1, 3
    return v;
}

// Sum over the lane's half-wave, in every lane of that half (the scan's last lanes hold the
// two sums).
__device__ __forceinline__ int half_total(int v) {
    const int inc = (int)half_scan((uint32_t)v);
    const int lo = __builtin_amdgcn_readlane(inc, 31), hi = __builtin_amdgcn_readlane(inc, 63);
    return (threadIdx.x & 32) ? hi : lo;
}

// Rows r0..r0+3 of mask m as bit 0 of bytes 0..3. The product is kept opaque so that a shift
// applied afterwards is not folded into the multiplier (a 24-bit multiply stays one).
__device__ __forceinline__ uint32_t stripe_bytes(uint32_t m, int r0) {
    uint32_t t = __umul24((m >> r0) & 0xFu, 0x204081u) & 0x01010101u;
    asm("" : "+v"(t));
    return t;
}

constexpr int kChunkBytes = 65536;  // symbol pool grabbed per wave at a time

}  // namespace

// ---------------------------------------------------------------- counters
__global__ void clear_counters_kernel(JobCounters* c, uint32_t* work, int nwork) {
    const int t = threadIdx.x;
    if (t < (int)(sizeof(JobCounters) / 4)) ((uint32_t*)c)[t] = 0;
    if (t < nwork) work[t] = 0;
}

hipError_t launch_clear_counters(JobCounters* c, uint32_t* work, int nwork, hipStream_t s) {
    hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, s, c, work, nwork);
    return hipGetLastError();
}

// ---------------------------------------------------------------- debug quantiser
// The coefficient words the context-modelling kernel quantises for itself, materialised
// (1,024 per block, row-major, zero-padded) for the debug taps.
__global__ void __launch_bounds__(256) quant_debug_kernel(const T1Block* blocks, const T1Band* bands,
                                                          const T1FrameRef* frames, uint32_t n, uint32_t* out) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t cb = (uint32_t)(gid >> 10);
    if (cb >= n) return;
    const int k = (int)(gid & 1023), r = k >> 5, x = k & 31;
    const T1Block b = blocks[cb];
    uint32_t word = 0;
    if (x < b.w && r < b.h) {
        const T1Band band = bands[b.band];
        const T1FrameRef fr = frames[b.frame];
        const float v = fr.coef[b.coef_off + (size_t)r * fr.stride + x];
        const float a = (v < 0.f ? -v : v) * band.inv_delta * 64.0f;
        const uint32_t q = a >= (float)band.max_mag ? band.max_mag : (uint32_t)a;
        word = q | ((v < 0.f && q >= 64u) ? 0x80000000u : 0u);
    }
    out[gid] = word;
}

hipError_t launch_quant_debug(const T1Block* blocks, const T1Band* bands, const T1FrameRef* frames, uint32_t n,
                              uint32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(quant_debug_kernel, dim3((unsigned)(((size_t)n * 1024 + 255) / 256)), dim3(256), 0, s, blocks,
                       bands, frames, n, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- context modelling
template <int WAVES>
__global__ void __launch_bounds__(WAVES * 64, 3) t1_cm_kernel(CmParams P) {
    // a plane's symbols per block, then 128 bytes where each lane writes the symbols its rows
    // do not have (so that every lane stores on every step)
    __shared__ __attribute__((aligned(16))) uint8_t s_sym[WAVES][2][kPlaneSymCap + 128];
    // Early-termination counts per context of the current plane (d = 0 low half, d = 1 high
    // half) and the contexts' entropy terms.
    __shared__ uint32_t s_hist[WAVES][2][32];
    __shared__ __attribute__((aligned(16))) float s_term[WAVES][2][20];

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, half = lane >> 5, x = lane & 31;
    uint8_t* const sym = s_sym[wv][half];
    uint8_t* const junk = sym + kPlaneSymCap + 4 * x;
    uint32_t* const hist = s_hist[wv][half];
    float* const terms = s_term[wv][half];
    unsigned long long chunk_cur = 0, chunk_end = 0;  // wave-uniform sub-allocator

    for (;;) {
        uint32_t pair = 0;
        if (lane == 0) pair = atomicAdd(P.work_next, 1u);
        pair = __shfl((int)pair, 0);
        if (2 * pair >= P.nblocks) break;
        const uint32_t cb = 2 * pair + half;
        const bool has = cb < P.nblocks;
        T1Block blk{};
        if (has) blk = P.blocks[cb];
        const int w = has ? blk.w : 0, h = has ? blk.h : 0;
        const T1Band band = P.bands[has ? blk.band : 0];
        const T1FrameRef fr = P.frames[has ? blk.frame : 0];
        const bool hl = has && blk.orient == 1, hh = has && blk.orient == 3;

        // Fused dead-zone quantisation of the block's column (quant.h:20-25).
        uint32_t m[32];
        uint32_t sgn = 0, colmax = 0;
        {
            // all 32 loads in flight at once: rows and columns outside the block read the
            // block's first sample (a valid address) and are discarded
            const float* src = fr.coef + blk.coef_off + (x < w ? x : 0);
            float v[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) v[r] = src[(size_t)(r < h ? r : 0) * fr.stride];
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const float a = (v[r] < 0.f ? -v[r] : v[r]) * band.inv_delta * 64.0f;
                uint32_t q = a >= (float)band.max_mag ? band.max_mag : (uint32_t)a;
                q = x < w && r < h ? q : 0u;
                sgn |= (v[r] < 0.f && q >= 64u ? 1u : 0u) << r;
                m[r] = q;
                colmax = q > colmax ? q : colmax;
            }
        }
        const uint32_t bmax = half_max_u(colmax) >> 6;
        const int nbps = bmax ? 32 - __clz((int)bmax) : 0;
        // the pair's plane and stripe counts, in scalar registers (wave-uniform loop bounds)
        const int nbps_lo = __builtin_amdgcn_readlane(nbps, 0), nbps_hi = __builtin_amdgcn_readlane(nbps, 32);
        const int jmax = nbps_lo > nbps_hi ? nbps_lo : nbps_hi;
        const int h_lo = __builtin_amdgcn_readlane(h, 0), h_hi = __builtin_amdgcn_readlane(h, 32);
        const int nstripes = ((h_lo > h_hi ? h_lo : h_hi) + 3) >> 2;
        const uint32_t sgnL = from_left(sgn, x), sgnR = from_right(sgn, x);
        const uint32_t sgnU = sgn << 1, sgnD = sgn >> 1;
        const uint32_t rowsValid = h >= 32 ? 0xFFFFFFFFu : ((1u << h) - 1u);
        const float thr = fr.slope_thresh;
        const bool track = thr > 0.f;
        if (x < 20) hist[x] = 0;

        // Bit-plane masks of the column: pl[k] holds bit (p + k) of every row's magnitude
        // (bits p..p+6: the current magnitude bit is pl[6], the six below it feed the
        // distortion estimates). Moving to the next plane shifts the register down by one
        // and extracts a single new mask.
        uint32_t pl[7];
        {
            const int p0 = nbps > 0 ? nbps - 1 : 0;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                uint32_t mk = 0;
#pragma unroll
                for (int r = 0; r < 32; ++r) mk |= __builtin_amdgcn_ubfe(m[r], (uint32_t)(p0 + k), 1u) << r;
                pl[k] = mk;
            }
        }
        // Sum of the 7-bit windows (bits p..p+6) of the rows in mask s.
        auto window_sum = [&](uint32_t s) {
            int t = 0;
#pragma unroll
            for (int k = 0; k < 7; ++k) t += __popc(pl[k] & s) << k;
            return t;
        };

        uint32_t S = 0, R = 0;  // significant, refined before
        uint32_t np = 0, sym_total = 0, planes = 0;
        bool done = !has || nbps == 0 || nbps > kMaxPlanes;
        for (int j = 0; j < jmax; ++j) {
            const bool act = !done && j < nbps;
            const int p = nbps - 1 - j;  // current bit-plane (valid when act)
            if (j > 0) {
#pragma unroll
                for (int k = 6; k > 0; --k) pl[k] = pl[k - 1];
                const int sh = p > 0 ? p : 0;
                uint32_t mk = 0;
#pragma unroll
                for (int r = 0; r < 32; ++r) mk |= __builtin_amdgcn_ubfe(m[r], (uint32_t)sh, 1u) << r;
                pl[0] = act ? mk : 0u;
            }
            const uint32_t Bp = pl[6];
            const uint32_t sig0 = S;
            uint32_t V = 0, plane_count = 0;
            int plane_dist = 0;
            for (int type = (j == 0 ? 2 : 0); type < 3; ++type) {
                const bool pact = act && np < P.max_passes;
                const uint32_t colMask = (pact && x < w) ? rowsValid : 0u;
                const uint32_t SL = from_left(S, x), SR = from_right(S, x);
                // The pass over the whole column: keep0 the rows whose zero-coding (or
                // refinement) symbol is written, keep1 those whose sign symbol is written,
                // rl the run-length symbols of each stripe (a nibble per stripe); m0 / m1 the
                // bits of those symbol bytes, one mask per bit.
                uint32_t keep0 = 0, keep1 = 0, runs = 0, rl = 0, add0 = 0;
                uint32_t m0[5] = {Bp, 0, 0, 0, 0}, m1[4] = {0, 0, 0, 0};
                int dist = 0;
                if (type == 1) {
                    // ---- magnitude refinement (D.3.3): samples significant before this plane;
                    // symbol 28 + (bit | (context - 14) << 1), context 16 if refined before,
                    // else 15 if a neighbour is significant, else 14
                    keep0 = sig0 & colMask;
                    m0[1] = neighbour_any(S, SL, SR) & ~R;
                    m0[2] = R;
                    add0 = 0x1C1C1C1Cu;
                    // t1_dist_ref summed: 64 w - 5120 where bit p+6 is set, 3072 - 64 w elsewhere
                    const uint32_t s1 = keep0 & Bp, s0 = keep0 & ~Bp;
                    dist = 64 * window_sum(s1) - 5120 * __popc(s1) + 3072 * __popc(s0) - 64 * window_sum(s0);
                    R |= keep0;
                } else {
                    uint32_t N = 0;  // samples the pass turns significant
                    if (type == 0) {
                        // ---- significance propagation (D.3.1): the least fixed point of "an
                        // insignificant row is coded when a neighbour is significant as the scan
                        // sees it"; each round every lane fills its column given the other
                        // columns' current new significance, until no lane's set changes
                        const uint32_t open = ~S & colMask;
                        const uint32_t around = neighbour_any(S, SL, SR);
                        uint32_t E = 0;
                        if (__any((open & around) != 0)) {
                            for (;;) {
                                const uint32_t seen = around | new_seen(from_left(N, x), from_right(N, x));
                                E = column_fill(open & seen, open, Bp);
                                const uint32_t turned = E & Bp;
                                if (!__any(turned != N)) break;
                                N = turned;
                            }
                        }
                        keep0 = E;
                        V |= E;
                    } else {
                        // ---- cleanup with run-length mode (D.3.4): every remaining sample
                        const uint32_t Cc = colMask & ~S & ~V;
                        N = Cc & Bp;
                        runs = run_stripes(Cc, S, SL, SR, N, from_left(N, x), from_right(N, x));
                        cleanup_run_symbols(Cc, N, runs, keep0, rl);
                    }
                    keep1 = N;
                    if (__any((keep0 | keep1 | rl) != 0)) {
                        const Seen v = seen_at_coding(S, SL, SR, N, from_left(N, x), from_right(N, x));
                        const ZcCtx z = zc_context(v, hl, hh);
                        const ScCtx g = sign_context(v, sgnL, sgnR, sgn, sgnU, sgnD);
                        m0[1] = z.c0;
                        m0[2] = z.c1;
                        m0[3] = z.c2;
                        m0[4] = z.c3;
                        m1[0] = g.d;
                        m1[1] = g.c0;
                        m1[2] = g.c1;
                        m1[3] = g.c2;
                    }
                    S |= N;
                    // t1_dist_sig summed over the new significances: 192 w - 9216 each
                    dist = 192 * window_sum(N) - 9216 * __popc(N);
                }
                if (__any((keep0 | keep1 | rl) != 0)) {
                    // symbols per stripe (a nibble each), placed stripe after stripe and, within
                    // a stripe, column after column by an exclusive prefix over the half-wave;
                    // stripes k and k + 4 share a register, 16 bits each
                    const uint32_t cnt = nibble_counts(keep0) + nibble_counts(keep1) + rl;
                    uint32_t e0, e1, e2, e3, t0, t1, t2, t3, u0, u1, u2, u3;
                    {
                        uint32_t f = cnt & 0x000F000Fu, inc = half_scan(f);
                        e0 = inc - f;
                        t0 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 31);
                        u0 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
                        f = (cnt >> 4) & 0x000F000Fu;
                        inc = half_scan(f);
                        e1 = inc - f;
                        t1 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 31);
                        u1 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
                        f = (cnt >> 8) & 0x000F000Fu;
                        inc = half_scan(f);
                        e2 = inc - f;
                        t2 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 31);
                        u2 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
                        f = (cnt >> 12) & 0x000F000Fu;
                        inc = half_scan(f);
                        e3 = inc - f;
                        t3 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 31);
                        u3 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
                    }
                    uint32_t base = plane_count;
                    for (int s = 0; s < nstripes; ++s) {
                        // this stripe's prefix and half-wave totals, then move the next ones up
                        const uint32_t ex = e0 & 0xFFFFu, tlo = t0 & 0xFFFFu, thi = u0 & 0xFFFFu;
                        {
                            const uint32_t a = e0 >> 16, b = t0 >> 16, c = u0 >> 16;
                            e0 = e1; e1 = e2; e2 = e3; e3 = a;
                            t0 = t1; t1 = t2; t2 = t3; t3 = b;
                            u0 = u1; u1 = u2; u2 = u3; u3 = c;
                        }
                        if (tlo + thi == 0) continue;
                        uint32_t pos = base + ex;
                        base += half ? thi : tlo;
                        const int r0 = 4 * s;
                        const uint32_t z = (keep0 >> r0) & 0xFu, gn = (keep1 >> r0) & 0xFu;
                        const uint32_t w0 = add0 + (stripe_bytes(m0[0], r0) | stripe_bytes(m0[1], r0) << 1 |
                                                    stripe_bytes(m0[2], r0) << 2 | stripe_bytes(m0[3], r0) << 3 |
                                                    stripe_bytes(m0[4], r0) << 4);
                        if (type == 2 && __any(((runs >> r0) & 1u) != 0)) {
                            // run-length symbols: context 17 with whether a row turns
                            // significant, then that row's position in two context-18 bits
                            const bool run = (runs >> r0) & 1u, hit = run && gn != 0;
                            const uint32_t first = gn ? (uint32_t)__builtin_ctz(gn) : 0u;
                            *(run ? sym + pos : junk) = (uint8_t)(34u | (gn != 0));
                            *(hit ? sym + pos + 1 : junk) = (uint8_t)(36u | (first >> 1));
                            *(hit ? sym + pos + 2 : junk) = (uint8_t)(36u | (first & 1u));
                            pos += run ? (hit ? 3u : 1u) : 0u;
                        }
                        if (type == 1) {
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const uint32_t zk = (z >> k) & 1u;
                                *(zk ? sym + pos : junk) = (uint8_t)(w0 >> (8 * k));
                                pos += zk;
                            }
                        } else {
                            const uint32_t w1 = 0x12121212u + (stripe_bytes(m1[0], r0) | stripe_bytes(m1[1], r0) << 1 |
                                                               stripe_bytes(m1[2], r0) << 2 | stripe_bytes(m1[3], r0) << 3);
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const uint32_t zk = (z >> k) & 1u, gk = (gn >> k) & 1u;
                                *(zk ? sym + pos : junk) = (uint8_t)(w0 >> (8 * k));
                                pos += zk;
                                *(gk ? sym + pos : junk) = (uint8_t)(w1 >> (8 * k));
                                pos += gk;
                            }
                        }
                    }
                    plane_count = base;
                }
                const int total = half_total(dist);
                if (pact) {
                    if (x == 0) {
                        P.pass_end[(size_t)cb * kMaxPasses + np] = sym_total + plane_count;
                        P.pass_dist[(size_t)cb * kMaxPasses + np] = total;
                    }
                    ++np;
                    plane_dist += total;
                }
            }
            // ---- the plane's symbols to the pool (wave-uniform sub-allocation)
            wave_sync();
            // early-termination counts of the plane's symbols, 32 lanes per block at a time
            // (not one lane per column inside the divergent emission loops)
            if (track && act)
                for (uint32_t i = 4u * (uint32_t)x; i < plane_count; i += 128u) {
                    const uint32_t w4 = *(const uint32_t*)(sym + i);  // within the staging area
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j)
                        if (i + j < plane_count) {
                            const uint32_t sv = (w4 >> (8 * j)) & 0xFFu;
                            atomicAdd(&hist[sv >> 1], (sv & 1u) ? 0x10000u : 1u);
                        }
                }
            const uint32_t cnt_lo = __shfl((int)plane_count, 0), cnt_hi = __shfl((int)plane_count, 32);
            // 64-byte aligned: the MQ coder reads a plane in 64-byte groups
            const uint32_t al_lo = (cnt_lo + 63u) & ~63u, al_hi = (cnt_hi + 63u) & ~63u;
            if (chunk_cur + al_lo + al_hi > chunk_end) {
                const unsigned long long want = kChunkBytes > al_lo + al_hi ? kChunkBytes : al_lo + al_hi;
                unsigned long long got = 0;
                if (lane == 0) got = atomicAdd(&P.counters->sym_next, want);
                got = __shfl((long long)got, 0);
                chunk_cur = got;
                chunk_end = got + want;
                if (chunk_end > P.pool_cap && lane == 0) atomicOr(&P.counters->flags, 1u);
            }
            const unsigned long long off = chunk_cur + (half ? al_lo : 0u);
            chunk_cur += al_lo + al_hi;
            const bool fits = chunk_end <= P.pool_cap;
            if (act && fits) {
                for (uint32_t i = (uint32_t)x * 16u; i < plane_count; i += 512u)
                    *(uint4*)(P.pool + off + i) = *(const uint4*)(sym + i);
                if (x == 0) P.chunks[(size_t)cb * kMaxPlanes + j] = make_uint2((uint32_t)(off >> 4), plane_count);
            }
            if (act) {
                sym_total += plane_count;
                ++planes;
            }
            wave_sync();
            // ---- early termination after plane p > 0 (t1_ref.cpp:128-162)
            bool stop_now = false;
            if (__any(track)) {
                float term = 0.f;
                if (track && x < 19) {
                    const uint32_t hv = hist[x];
                    const uint32_t n0 = hv & 0xFFFFu, n1 = hv >> 16, n = n0 + n1;
                    if (n > 0 && n <= kLog2Max) {
                        const float* row = P.log2tab + (size_t)n * (n + 1) / 2;
                        const float a = (float)n0 * row[n0];
                        const float b = (float)n1 * row[n1];
                        term = -(a + b) + 2.f;
                    }
                }
                if (x < 20) terms[x] = term;
                wave_sync();
                if (track && x < 20) hist[x] = 0;
                float bits = 0.f;  // the contexts' terms summed in context order
                {
                    float4 tv[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) tv[k] = ((const float4*)terms)[k];
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        bits += tv[k].x;
                        bits += tv[k].y;
                        bits += tv[k].z;
                        if (k < 4) bits += tv[k].w;
                    }
                }
                const float bytes = bits * 0.125f + 1.f;
                bool cand = track && act && p > 0;
                float dabs = 0.f;
                if (cand) dabs = (float)plane_dist * ldexpf(1.0f, 2 * p) * (1.f / 4096.f) * band.dist_scale;
                cand = cand && dabs < thr * bytes;
                if (__any(cand)) {
                    // Remaining squared error (t1_ref.cpp:142-154): the CPU sums the samples'
                    // float terms in row-major order. That float sum of at most 1,024
                    // non-negative terms is within 1023 * 2^-24 (< 6.2e-5, relative) of their
                    // exact sum; the half-wave's sum (float column sums, within 31 * 2^-24,
                    // then a double shuffle tree) is within 2e-6 of it. Unless that sum lies
                    // within 1e-4 of the threshold it decides the comparison the same way;
                    // only then is the row-major float sum formed.
                    const int pp = act ? p : 0;
                    const float step = ldexpf(1.0f, pp);
                    auto term = [&](uint32_t mm) {
                        const float v = (float)mm * (1.f / 64.f);
                        const uint32_t q = mm >> (pp + 6);
                        const float rec = q == 0 ? 0.f : ((float)q + 0.5f) * step;
                        const float dd = v - rec;
                        return dd * dd;
                    };
                    float col = 0.f;
#pragma unroll
                    for (int r = 0; r < 32; ++r)
                        if (x < w && r < h) col += term(m[r]);
                    const double lhs = half_sum_d((double)col) * (double)band.dist_scale;
                    const double T = (double)(thr * bytes);
                    const bool low = lhs * (1.0 + 1e-4) < T, high = lhs * (1.0 - 1e-4) >= T;
                    const bool unsure = cand && !low && !high;
                    bool stop = cand && low;
                    if (__any(unsure)) {
                        float e = 0.f;
                        float* scratch = (float*)sym;
#pragma unroll
                        for (int part = 0; part < 2; ++part) {
#pragma unroll
                            for (int rr = 0; rr < 16; ++rr) {
                                uint32_t mm = m[part * 16 + rr];
                                asm volatile("" : "+v"(mm));  // recomputed here, not kept from the fast path
                                scratch[rr * 32 + x] = term(mm);
                            }
                            wave_sync();
                            if (x == 0 && unsure) {
                                const int r1 = h < part * 16 + 16 ? h : part * 16 + 16;
                                for (int r = part * 16; r < r1; ++r)
                                    for (int c = 0; c < w; ++c) e += scratch[(r - part * 16) * 32 + c];
                            }
                            wave_sync();
                        }
                        e = __shfl(e, half << 5);
                        if (unsure) stop = e * band.dist_scale < thr * bytes;
                    }
                    stop_now = stop;
                }
            }
            if (act && (stop_now || np >= P.max_passes)) done = true;
            if (!act) done = true;
        }
        if (has && x == 0) {
            T1Info inf;
            inf.nbps = (uint32_t)(nbps <= kMaxPlanes ? nbps : 0);
            inf.passes = np;
            inf.planes = planes;
            inf.symbols = sym_total;
            P.info[cb] = inf;
            if (nbps > kMaxPlanes) atomicOr(&P.counters->flags, 4u);
        }
    }
}

hipError_t launch_t1_cm(const CmParams& p, int waves, int groups, hipStream_t s) {
    switch (waves) {
    case 1: hipLaunchKernelGGL(t1_cm_kernel<1>, dim3(groups), dim3(64), 0, s, p); break;
    case 2: hipLaunchKernelGGL(t1_cm_kernel<2>, dim3(groups), dim3(128), 0, s, p); break;
    case 8: hipLaunchKernelGGL(t1_cm_kernel<8>, dim3(groups), dim3(512), 0, s, p); break;
    default: hipLaunchKernelGGL(t1_cm_kernel<4>, dim3(groups), dim3(256), 0, s, p); break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- ordering
// Counting sort of the blocks by symbol count, largest first: 128 buckets, four per
// octave. Kernel 1 histograms, kernel 2 turns the histogram into bucket starts, kernel 3
// scatters (one global reservation per bucket and workgroup).
namespace {
constexpr int kBuckets = 128;
constexpr int kSortSpan = 4096;  // blocks per workgroup
__device__ __forceinline__ uint32_t bucket_of(uint32_t s) {
    if (s < 4) return kBuckets - 1;
    const int l = 31 - __clz((int)s);
    const uint32_t sub = (s >> (l - 2)) & 3u;
    const uint32_t b = (uint32_t)l * 4 + sub;
    return kBuckets - 1 - (b < kBuckets - 1 ? b : kBuckets - 1);
}
}  // namespace

__global__ void __launch_bounds__(256) sort_hist_kernel(const T1Info* info, uint32_t n, uint32_t* hist,
                                                        JobCounters* counters) {
    __shared__ uint32_t h[kBuckets];
    __shared__ unsigned long long total;
    for (int i = threadIdx.x; i < kBuckets; i += 256) h[i] = 0;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kSortSpan, b1 = min(n, b0 + kSortSpan);
    unsigned long long mine = 0;
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256) {
        const uint32_t s = info[i].symbols;
        mine += s;
        atomicAdd(&h[bucket_of(s)], 1u);
    }
    atomicAdd(&total, mine);
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets; i += 256)
        if (h[i]) atomicAdd(&hist[i], h[i]);
    if (threadIdx.x == 0) atomicAdd(&counters->symbols, total);
}

__global__ void sort_scan_kernel(uint32_t* hist) {
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 0; b < kBuckets; ++b) {
            const uint32_t c = hist[b];
            hist[b] = run;
            run += c;
        }
    }
}

__global__ void __launch_bounds__(256) sort_scatter_kernel(const T1Info* info, uint32_t n, uint32_t* cursor,
                                                           uint32_t* perm) {
    __shared__ uint32_t h[kBuckets], base[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kSortSpan, b1 = min(n, b0 + kSortSpan);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256) atomicAdd(&h[bucket_of(info[i].symbols)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets; i += 256) {
        base[i] = h[i] ? atomicAdd(&cursor[i], h[i]) : 0u;
        h[i] = 0;
    }
    __syncthreads();
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256) {
        const uint32_t b = bucket_of(info[i].symbols);
        perm[base[b] + atomicAdd(&h[b], 1u)] = i;
    }
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t launch_sort_blocks(const T1Info* info, uint32_t n, uint32_t* hist, uint32_t* perm, JobCounters* counters,
                              hipStream_t s) {
    const uint32_t groups = (n + kSortSpan - 1) / kSortSpan;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(128), 0, s, hist, kBuckets);
    hipLaunchKernelGGL(sort_hist_kernel, dim3(groups), dim3(256), 0, s, info, n, hist, counters);
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(64), 0, s, hist);
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(groups), dim3(256), 0, s, info, n, hist, perm);
    return hipGetLastError();
}

// ---------------------------------------------------------------- exclusive scans
// Three-phase exclusive scan over per-block values (kScanSpan blocks per workgroup):
// per-group totals, a one-group scan of the totals, then local scans plus the group base.
namespace {
constexpr int kScanSpan = 2048;

__device__ __forceinline__ uint32_t out_capacity(const T1Info& f) {
    // A block's codeword never needs more than half a byte per symbol plus the flush; a
    // block that still overflows is reported and the job re-run with more room.
    return f.passes ? ((f.symbols / 2u + 64u + 63u) & ~63u) : 0u;
}

template <class F>
__device__ void block_scan_u64(uint32_t n, unsigned long long* out, unsigned long long* partial, int phase, F value) {
    __shared__ unsigned long long s[256];
    const uint32_t b0 = blockIdx.x * kScanSpan, b1 = min(n, b0 + kScanSpan);
    const uint32_t per = kScanSpan / 256;
    const uint32_t i0 = b0 + threadIdx.x * per, i1 = min(b1, i0 + per);
    unsigned long long sum = 0;
    for (uint32_t i = i0; i < i1; ++i) sum += value(i);
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const unsigned long long v = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0ull;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (phase == 0) {
        if (threadIdx.x == 255) partial[blockIdx.x] = s[255];
        return;
    }
    unsigned long long run = partial[blockIdx.x] + s[threadIdx.x] - sum;
    for (uint32_t i = i0; i < i1; ++i) {
        out[i] = run;
        run += value(i);
    }
}
}  // namespace

__global__ void __launch_bounds__(256) out_offsets_kernel(const T1Info* info, uint32_t n, unsigned long long* out,
                                                          unsigned long long* partial, int phase) {
    block_scan_u64(n, out, partial, phase, [&](uint32_t i) { return (unsigned long long)out_capacity(info[i]); });
}

__global__ void scan_partials_kernel(unsigned long long* partial, int groups, unsigned long long* total) {
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int g = 0; g < groups; ++g) {
            const unsigned long long c = partial[g];
            partial[g] = run;
            run += c;
        }
        if (total) *total = run;
    }
}

hipError_t launch_out_offsets(const T1Info* info, uint32_t n, unsigned long long* out_off, unsigned long long* partial,
                              hipStream_t s) {
    const int groups = (int)((n + kScanSpan - 1) / kScanSpan);
    hipLaunchKernelGGL(out_offsets_kernel, dim3(groups), dim3(256), 0, s, info, n, out_off, partial, 0);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(64), 0, s, partial, groups, partial + groups);
    hipLaunchKernelGGL(out_offsets_kernel, dim3(groups), dim3(256), 0, s, info, n, out_off, partial, 1);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) data_offsets_kernel(const T1Info* info, const uint32_t* len, uint32_t n,
                                                           unsigned long long* out, unsigned long long* partial,
                                                           int phase) {
    block_scan_u64(n, out, partial, phase, [&](uint32_t i) {
        return (unsigned long long)((len[i] + 3u) & ~3u) | ((unsigned long long)info[i].passes << 36);
    });
}

hipError_t launch_compact_offsets(const T1Info* info, const uint32_t* code_len, uint32_t n,
                                  unsigned long long* data_off, uint32_t* pass_off, unsigned long long* partial,
                                  hipStream_t s) {
    (void)pass_off;
    const int groups = (int)((n + kScanSpan - 1) / kScanSpan);
    hipLaunchKernelGGL(data_offsets_kernel, dim3(groups), dim3(256), 0, s, info, code_len, n, data_off, partial, 0);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(64), 0, s, partial, groups, partial + groups);
    hipLaunchKernelGGL(data_offsets_kernel, dim3(groups), dim3(256), 0, s, info, code_len, n, data_off, partial, 1);
    return hipGetLastError();
}

// ---------------------------------------------------------------- MQ coding
namespace {
// Table C.2 packed as Qe | NMPS << 16 | NLPS << 22 | SWITCH << 28. A context's state word is
// its entry of this table with its MPS in bit 31.
__constant__ uint32_t kMqTable[47];

}  // namespace

// Truncation length of every pass but the last (mq_encoder.h:146-169): the shortest prefix,
// never ending in 0xFF, that with the decoder's 1-fill stays inside the interval the coder had
// at the pass end. out holds the block's len codeword bytes.
__device__ __forceinline__ void mq_pass_lengths(const uint8_t* out, uint32_t len, const MqSnapshot* snaps,
                                                uint32_t np, uint32_t* plen) {
    uint32_t prev = 1;
    for (uint32_t q = 0; q + 1 < np; ++q) {
        const MqSnapshot sn = snaps[q];
        const int ct = (int)(sn.b_ct >> 16);
        const int base_sh = 43 - ct;
        const unsigned long long low =
            ((unsigned long long)(sn.b_ct & 0xFFFFu) << base_sh) + ((unsigned long long)sn.c << 16);
        const unsigned long long high = low + ((unsigned long long)sn.a << 16);
        unsigned long long prefix = 0;
        int rel = 0;
        uint32_t L = len;
        for (int i = sn.bp; i < (int)len; ++i) {
            const uint32_t byte = i >= 0 ? out[i] : 0u;
            const int sh = base_sh - rel;
            if (sh < 0) break;
            prefix += (unsigned long long)byte << sh;
            const unsigned long long tail = prefix + (1ull << sh);
            if ((uint32_t)(i + 1) >= prev && byte != 0xFFu && tail <= high && tail > low) {
                L = (uint32_t)(i + 1);
                break;
            }
            rel += byte == 0xFFu ? 7 : 8;
        }
        plen[q] = L;
        prev = L > prev ? L : prev;
    }
    plen[np - 1] = len;
}

// One code-block coded by a whole wave, for the longest blocks of a launch: a lane-per-block
// wave advances all its blocks one symbol per step of about a hundred instructions, so its
// longest block sets the launch's time. Here the coder's registers are wave-uniform (scalar)
// values, the 19 context states sit in lanes 0-18 of one vector register and Table C.2 in
// lanes 0-46 of another (tabv), both read with readlane, and a symbol is a short chain of
// scalar instructions. The bytes are those of the lane coder (the same steps on the same
// state). There is no lane-dependent branch: every lane runs the same code and stores the same
// values to the same addresses, so the compiler can never run the wave's lanes apart (with a
// lane-0-only tail it had sent the other lanes into the next block's loop without lane 0).
__device__ void mq_code_wave(const MqParams& P, uint32_t cb, uint32_t tabv, int lane) {
    const T1Info inf = P.info[cb];
    const uint32_t np = inf.passes;
    if (np == 0) {
        P.code_len[cb] = 0;
        return;
    }
    const unsigned long long obase = P.out_off[cb];
    uint8_t* const out = P.out + obase;
    uint32_t cap = out_capacity(inf);  // a multiple of 64, as every block's offset
    bool overflow = obase + cap > P.out_cap;
    if (overflow) cap = 0;
    const uint32_t q0 = (uint32_t)__builtin_amdgcn_readlane((int)tabv, 0);
    const uint32_t q3 = (uint32_t)__builtin_amdgcn_readlane((int)tabv, 3);
    const uint32_t q4 = (uint32_t)__builtin_amdgcn_readlane((int)tabv, 4);
    const uint32_t q46 = (uint32_t)__builtin_amdgcn_readlane((int)tabv, 46);
    uint32_t ctxv = q0;
    ctxv = lane == 0 ? q4 : ctxv;
    ctxv = lane == 17 ? q3 : ctxv;
    ctxv = lane == 18 ? q46 : ctxv;
    uint32_t A = 0x8000u, C = 0u, B = 0u;
    int CT = 12, bp = -1;
    unsigned long long acc = 0ull;  // finished bytes of the current aligned 8-byte group
    uint32_t flushed = 0;
    auto byteout = [&]() {
        const uint32_t carry = B != 0xFFu ? (C >> 27) & 1u : 0u;
        B += carry;
        C -= carry << 27;
        const bool stuff = B == 0xFFu;
        acc |= (unsigned long long)B << (8 * (bp & 7));
        if (bp >= 0 && (bp & 7) == 7) {
            const uint32_t at = (uint32_t)bp & ~7u;
            if (at < cap) *(unsigned long long*)(out + at) = acc;
            else overflow = true;
            acc = 0ull;
            flushed = at + 8u;
        }
        ++bp;
        B = stuff ? C >> 20 : (C >> 19) & 0xFFu;
        C &= stuff ? 0xFFFFFu : 0x7FFFFu;
        CT = stuff ? 7 : 8;
    };
    const uint32_t* pend = P.pass_end + (size_t)cb * kMaxPasses;
    MqSnapshot* snaps = P.snaps + (size_t)cb * kMaxPasses;
    uint32_t k = 0, pos = 0, next_end = pend[0];
    const uint2* chunk = P.chunks + (size_t)cb * kMaxPlanes;
    for (uint32_t j = 0; j < inf.planes; ++j) {
        const uint2 ch = chunk[j];
        const uint8_t* src = P.pool + ((unsigned long long)ch.x << 4);
        for (uint32_t base = 0; base < ch.y; base += 256u) {
            // 256 symbols at a time, 4 per lane; lanes past the plane's end read its first
            // word again (a plane's symbols start 64-byte aligned)
            const uint32_t cnt = ch.y - base < 256u ? ch.y - base : 256u;
            const uint32_t lo = 4u * (uint32_t)lane < cnt ? 4u * (uint32_t)lane : 0u;
            const uint32_t wv = *(const uint32_t*)(src + base + lo);
            for (uint32_t i = 0; i < cnt; ++i) {
                while (pos == next_end) {  // pass ends: the truncation reference of the pass
                    snaps[k] = MqSnapshot{bp, B | ((uint32_t)CT << 16), C, A};
                    ++k;
                    next_end = k < np ? pend[k] : 0xFFFFFFFFu;
                }
                const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)wv, (int)(i >> 2));
                const uint32_t sv = (word >> (8u * (i & 3u))) & 0xFFu;
                const uint32_t cx = sv >> 1;
                const uint32_t st = (uint32_t)__builtin_amdgcn_readlane((int)ctxv, (int)cx);
                const uint32_t qe = st & 0xFFFFu;
                const uint32_t a = A - qe;
                const bool is_mps = (int)(st ^ (sv << 31)) >= 0;
                const bool add_c = is_mps != (a < qe);
                C += add_c ? qe : 0u;
                A = add_c ? a : qe;
                if (!is_mps || a < 0x8000u) {
                    const uint32_t nidx = is_mps ? ((st >> 16) & 63u) : ((st >> 22) & 63u);
                    const uint32_t msb = is_mps ? st : st ^ (st << 3);
                    const uint32_t t_new =
                        (msb & 0x80000000u) | (uint32_t)__builtin_amdgcn_readlane((int)tabv, (int)nidx);
                    ctxv = lane == (int)cx ? t_new : ctxv;
                    int nsh = __clz((int)A) - 16;
                    A <<= nsh;
                    while (nsh >= CT) {
                        C <<= CT;
                        nsh -= CT;
                        byteout();
                    }
                    C <<= nsh;
                    CT -= nsh;
                }
                ++pos;
            }
        }
    }
    while (pos == next_end) {
        snaps[k] = MqSnapshot{bp, B | ((uint32_t)CT << 16), C, A};
        ++k;
        next_end = k < np ? pend[k] : 0xFFFFFFFFu;
    }
    // FLUSH (C.2.9)
    const uint32_t tempc = C + A;
    C |= 0xFFFFu;
    if (C >= tempc) C -= 0x8000u;
    C <<= CT;
    byteout();
    C <<= CT;
    byteout();
    uint32_t len = (uint32_t)bp;
    if (B != 0xFFu) {
        acc |= (unsigned long long)B << (8 * (bp & 7));
        len = (uint32_t)bp + 1;
    }
    if (len > cap) overflow = true;
    if (overflow) {
        atomicOr(&P.counters->flags, 2u);
        P.code_len[cb] = 0;
        return;
    }
    const uint32_t nw = (len + 3u) >> 2;
    for (uint32_t wi = flushed >> 2; wi < nw; ++wi)
        *(uint32_t*)(out + 4u * wi) = (uint32_t)(acc >> (32u * (wi & 1u)));
    mq_pass_lengths(out, len, snaps, np, P.pass_len + (size_t)cb * kMaxPasses);
    P.code_len[cb] = len;
}

// GROUP lanes per workgroup; NQ: 16-byte words of symbols read per lane at a time (1 or 4),
// with the following NQ in flight while these are coded.
template <int GROUP, int NQ>
__global__ void __launch_bounds__(GROUP) t1_mq_kernel(MqParams P) {
    __shared__ uint32_t s_ctx[19][GROUP];
    __shared__ uint32_t s_tab[48];
    if (threadIdx.x < 47) s_tab[threadIdx.x] = kMqTable[threadIdx.x];
    __syncthreads();
    // A full symbol pool left some chunk tables unwritten: nothing to code, the host
    // re-runs Tier-1 with a larger pool.
    if (P.counters->flags & 1u) return;
    constexpr uint32_t kGroupBytes = 16u * NQ;
    const int t = threadIdx.x;
    // The nlong longest blocks (the leading entries of the order) one per wave first, by a
    // fixed stride over the grid's waves (the grid has at least nlong waves when it can).
    {
        const int lane = t & 63;
        const uint32_t tabv = lane < 47 ? s_tab[lane] : 0u;
        const uint32_t wave =
            (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (GROUP / 64) + ((uint32_t)t >> 6)));
        const uint32_t nwaves = gridDim.x * (GROUP / 64);
        for (uint32_t i = wave; i < P.nlong; i += nwaves) mq_code_wave(P, P.perm[i], tabv, lane);
    }
    const uint32_t nrest = P.nblocks - P.nlong;
    for (;;) {
        // one wave-sized group of blocks at a time, in largest-first order
        uint32_t base = 0;
        if ((t & 63) == 0) base = atomicAdd(P.work_next, 64u);
        base = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)base, 0));
        if (base >= nrest) break;
        const uint32_t gid = P.nlong + base + (t & 63);
        if (gid >= P.nblocks) continue;
        const uint32_t cb = P.perm[gid];
        const T1Info inf = P.info[cb];
        const uint32_t np = inf.passes;
        if (np == 0) {
            P.code_len[cb] = 0;
            continue;
        }
        const unsigned long long obase = P.out_off[cb];
        uint8_t* out = P.out + obase;
        uint32_t cap = out_capacity(inf);  // a multiple of 64, as every block's offset
        uint32_t overflow = obase + cap > P.out_cap ? 1u : 0u;  // a word, not a lane mask, in the hot loop
        if (overflow) cap = 0;
        for (int i = 0; i < 19; ++i) s_ctx[i][t] = s_tab[0];
        s_ctx[0][t] = s_tab[4];
        s_ctx[17][t] = s_tab[3];
        s_ctx[18][t] = s_tab[46];
        uint32_t A = 0x8000u, C = 0u, B = 0u;
        int CT = 12, bp = -1;
        // Finished bytes of the current aligned 8-byte group. The byte at bp = -1 is the
        // coder's initial dummy: it is 0 and takes no carry (C + A stays within 2^27 until the
        // first byte-out), so OR-ing it in changes nothing.
        unsigned long long acc = 0ull;
        uint32_t flushed = 0;           // codeword bytes already written out (a multiple of 8)
        // BYTEOUT (C.2.7) with the pending byte B kept in a register: it is final once the
        // next byte starts. Branch-free but for the 8-byte flush; the wave runs it whenever
        // one lane needs it, so it is kept short.
        auto byteout = [&]() {
            // a carry into B unless B is 0xFF; a B of 0xFF (before or after the carry) is
            // followed by a stuffed byte of 7 bits
            const uint32_t carry = B != 0xFFu ? (C >> 27) & 1u : 0u;
            B += carry;
            C -= carry << 27;
            const bool stuff = B == 0xFFu;
            acc |= (unsigned long long)B << (8 * (bp & 7));
            if (((uint32_t)bp & 0x80000007u) == 7u) {  // bp >= 0 ends a group
                const uint32_t at = (uint32_t)bp & ~7u;
                if (at < cap) *(unsigned long long*)(out + at) = acc;  // cap: a multiple of 64
                else overflow = 1u;
                acc = 0ull;
                flushed = at + 8u;
            }
            ++bp;
            B = stuff ? C >> 20 : (C >> 19) & 0xFFu;
            C &= stuff ? 0xFFFFFu : 0x7FFFFu;
            CT = stuff ? 7 : 8;
        };
        const uint32_t* pend = P.pass_end + (size_t)cb * kMaxPasses;
        MqSnapshot* snaps = P.snaps + (size_t)cb * kMaxPasses;
        // the next pass end and the one after it: the load of each pass end is issued a
        // whole pass before the coder reaches it
        uint32_t k = 0, pos = 0, next_end = pend[0], after_end = np > 1 ? pend[1] : 0xFFFFFFFFu;
        const uint2* chunk = P.chunks + (size_t)cb * kMaxPlanes;
        const uint32_t planes = inf.planes;
        uint2 ch = chunk[0];
        // A plane's symbols start 64-byte aligned and take a multiple of 64 bytes of the
        // pool, so whole groups can be read.
        const uint8_t* src = P.pool + ((unsigned long long)ch.x << 4);
        uint4 q[NQ];
#pragma unroll
        for (int e = 0; e < NQ; ++e) q[e] = *(const uint4*)(src + 16 * e);
        // the state of the next symbol's context and that context's slot
        uint32_t* cslot = &s_ctx[(q[0].x & 0xFFu) >> 1][t];
        uint32_t t_cur = *cslot;
        for (uint32_t j = 0; j < planes; ++j) {
            const uint2 chn = j + 1 < planes ? chunk[j + 1] : make_uint2(0u, 0u);
            const uint8_t* srcn = P.pool + ((unsigned long long)chn.x << 4);
            const uint32_t ngroups = (ch.y + kGroupBytes - 1u) / kGroupBytes;
            if (ngroups == 0 && j + 1 < planes) {  // an empty plane: the next plane's group now
#pragma unroll
                for (int e = 0; e < NQ; ++e) q[e] = *(const uint4*)(srcn + 16 * e);
            }
            for (uint32_t g = 0; g < ngroups; ++g) {
                const uint8_t* nsrc = g + 1 < ngroups ? src + kGroupBytes * (g + 1) : (j + 1 < planes ? srcn : nullptr);
                uint4 n[NQ];
#pragma unroll
                for (int e = 0; e < NQ; ++e) n[e] = nsrc ? *(const uint4*)(nsrc + 16 * e) : make_uint4(0u, 0u, 0u, 0u);
                const uint32_t mg = ch.y - kGroupBytes * g < kGroupBytes ? ch.y - kGroupBytes * g : kGroupBytes;
                // one copy of the 16 unrolled symbol steps: the group's 16-byte words move
                // to the front of q in turn
#pragma unroll 1
                for (int sub = 0; sub < NQ; ++sub) {
                    if (16u * sub < mg) {
                        const uint32_t m = mg - 16u * sub < 16u ? mg - 16u * sub : 16u;
                        // the symbol after this run of m: in the group, or the first of the next
                        const uint32_t after = 16u * sub + 16u < mg ? (q[NQ > 1 ? 1 : 0].x & 0xFFu) : (n[0].x & 0xFFu);
                        const uint4 cur = q[0];
                        const uint32_t w[5] = {cur.x, cur.y, cur.z, cur.w, after};
#pragma unroll
                        for (uint32_t b = 0; b < 16; ++b) {
                            if (b < m) {  // a guard, not an exit, so the 16 steps unroll
                                // pass ends: the coder state is the truncation reference of the pass
                                while (pos == next_end) {
                                    snaps[k] = MqSnapshot{bp, B | ((uint32_t)CT << 16), C, A};
                                    ++k;
                                    next_end = after_end;
                                    after_end = k + 1 < np ? pend[k + 1] : 0xFFFFFFFFu;
                                }
                                const uint32_t sv = (w[b >> 2] >> (8 * (b & 3))) & 0xFFu;
                                const uint32_t sn = b + 1 < m ? (w[(b + 1) >> 2] >> (8 * ((b + 1) & 3))) & 0xFFu : w[4];
                                // the next symbol's context state, read before this symbol may change it
                                uint32_t* const nslot = &s_ctx[sn >> 1][t];
                                const uint32_t t_next = *nslot;
                                const uint32_t st = t_cur;
                                const uint32_t qe = st & 0xFFFFu;
                                const uint32_t a = A - qe;  // A <= 0xFFFF, so a < 0x10000
                                const bool is_mps = (int)(st ^ (sv << 31)) >= 0;  // the decision equals the MPS
                                const bool add_c = is_mps != (a < qe);  // conditional exchange (C.2.5 / C.2.6)
                                C += add_c ? qe : 0u;
                                A = add_c ? a : qe;
                                // The state changes and A, C renormalise unless the symbol is
                                // the MPS with A still >= 0x8000; then the shift count below is
                                // 0 and the state written back is the old one, so every lane
                                // runs the same instructions and only byte-out branches. After
                                // an LPS the MPS flips where SWITCH (bit 28) is set.
                                const bool renorm = !is_mps || a < 0x8000u;
                                const uint32_t nidx = is_mps ? ((st >> 16) & 63u) : ((st >> 22) & 63u);
                                const uint32_t msb = is_mps ? st : st ^ (st << 3);
                                const uint32_t t_new = (msb & 0x80000000u) | s_tab[nidx];
                                const uint32_t st_new = renorm ? t_new : st;
                                *cslot = st_new;
                                // RENORME: A and C shift until A >= 0x8000 (at most 15 bits), a
                                // byte out whenever CT runs out; a second one is rare (an LPS
                                // with a small Qe). A is never 0 here.
                                int nsh = __builtin_clz(A) - 16;
                                A <<= nsh;
                                int sh = nsh < CT ? nsh : CT;
                                C <<= sh;
                                CT -= sh;
                                if (CT == 0) {
                                    nsh -= sh;
                                    do {
                                        byteout();
                                        sh = nsh < CT ? nsh : CT;
                                        C <<= sh;
                                        CT -= sh;
                                        nsh -= sh;
                                    } while (CT == 0);
                                }
                                t_cur = nslot == cslot ? st_new : t_next;
                                cslot = nslot;
                                ++pos;
                            }
                        }
                    }
#pragma unroll
                    for (int e = 0; e + 1 < NQ; ++e) q[e] = q[e + 1];
                }
#pragma unroll
                for (int e = 0; e < NQ; ++e) q[e] = n[e];
            }
            ch = chn;
            src = srcn;
        }
        while (pos == next_end) {
            snaps[k] = MqSnapshot{bp, B | ((uint32_t)CT << 16), C, A};
            ++k;
            next_end = after_end;
            after_end = k + 1 < np ? pend[k + 1] : 0xFFFFFFFFu;
        }
        // FLUSH (C.2.9)
        const uint32_t tempc = C + A;
        C |= 0xFFFFu;
        if (C >= tempc) C -= 0x8000u;
        C <<= CT;
        byteout();
        C <<= CT;
        byteout();
        uint32_t len = (uint32_t)bp;
        if (B != 0xFFu) {
            acc |= (unsigned long long)B << (8 * (bp & 7));
            len = (uint32_t)bp + 1;
        }
        // the bytes after the last full group, as whole words
        if (len > cap) {
            overflow = 1u;
        } else {
            const uint32_t nw = (len + 3u) >> 2;
            for (uint32_t wi = flushed >> 2; wi < nw; ++wi)
                *(uint32_t*)(out + 4u * wi) = (uint32_t)(acc >> (32u * (wi & 1u)));
        }
        if (overflow) {
            atomicOr(&P.counters->flags, 2u);
            P.code_len[cb] = 0;
            continue;
        }
        mq_pass_lengths(out, len, snaps, np, P.pass_len + (size_t)cb * kMaxPasses);
        P.code_len[cb] = len;
    }
}

static hipError_t upload_mq_table() {
    uint32_t tab[47];
    build_mq_table(tab);
    return hipMemcpyToSymbol(HIP_SYMBOL(kMqTable), tab, sizeof tab);
}

hipError_t t1_tables_init() { return upload_mq_table(); }

hipError_t launch_t1_mq(const MqParams& p, int group, int read_bytes, int groups, hipStream_t s) {
    const bool wide = read_bytes >= 64;
    switch (group) {
    case 64:
        if (wide) hipLaunchKernelGGL((t1_mq_kernel<64, 4>), dim3(groups), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((t1_mq_kernel<64, 1>), dim3(groups), dim3(64), 0, s, p);
        break;
    case 128:
        if (wide) hipLaunchKernelGGL((t1_mq_kernel<128, 4>), dim3(groups), dim3(128), 0, s, p);
        else hipLaunchKernelGGL((t1_mq_kernel<128, 1>), dim3(groups), dim3(128), 0, s, p);
        break;
    default:
        if (wide) hipLaunchKernelGGL((t1_mq_kernel<256, 4>), dim3(groups), dim3(256), 0, s, p);
        else hipLaunchKernelGGL((t1_mq_kernel<256, 1>), dim3(groups), dim3(256), 0, s, p);
        break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- compaction
// One wave per block: info record, pass tables and the codeword into dense arrays.
__global__ void __launch_bounds__(256) compact_kernel(CompactParams P) {
    if (P.counters->flags & 3u) return;  // the run is repeated with larger buffers
    const uint32_t cb = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t l = threadIdx.x & 63;
    if (cb >= P.nblocks) return;
    const T1Info inf = P.info[cb];
    const uint32_t len = P.code_len[cb];
    const unsigned long long packed = P.data_off[cb];
    const unsigned long long doff = packed & ((1ull << 36) - 1);
    const uint32_t poff = (uint32_t)(packed >> 36);
    if (l == 0) {
        uint32_t* di = P.dense_info + 5 * (size_t)cb;
        di[0] = inf.nbps;
        di[1] = inf.passes;
        di[2] = (uint32_t)doff;
        di[3] = len;
        di[4] = poff;
    }
    if (l < inf.passes) {
        P.dense_pass_len[poff + l] = P.pass_len[(size_t)cb * kMaxPasses + l];
        P.dense_pass_dist[poff + l] = P.pass_dist[(size_t)cb * kMaxPasses + l];
    }
    const uint32_t words = (len + 3) >> 2;
    const uint32_t* s = (const uint32_t*)(P.out + P.out_off[cb]);
    uint32_t* d = (uint32_t*)(P.dense_data + doff);
    for (uint32_t i = l; i < words; i += 64) d[i] = s[i];
}

hipError_t launch_compact(const CompactParams& p, hipStream_t s) {
    hipLaunchKernelGGL(compact_kernel, dim3((p.nblocks + 3) / 4), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace gail_amd
