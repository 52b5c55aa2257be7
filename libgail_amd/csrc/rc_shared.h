// PCRD-opt rate control written once for the host and the GPU (rate_control.cpp:11-279,
// t2.cpp:31-115): the rate-distortion hull of a code-block, the exact Tier-2 size of one
// packet, and the lambda search as a state machine that asks for size measurements. The
// host library runs it sequentially; the device runs it with one workgroup per frame, the
// measurements spread over the workgroup. Both do every floating-point step in the same
// order (no contraction), so they make the same choices.
#pragma once
#include <cstdint>

#include "device_types.h"

#if defined(__HIPCC__)
#define GAIL_HD __host__ __device__ inline
#define GAIL_UNROLL _Pragma("unroll")
#else
#define GAIL_HD inline
#define GAIL_UNROLL _Pragma("GCC unroll 16")
#endif

namespace gail_amd {

GAIL_HD double rc_inf() { return __builtin_huge_val(); }
GAIL_HD bool rc_finite(double v) { return v == v && v != rc_inf() && v != -rc_inf(); }

// Distortion scale of pass k of a block with nbps bit-planes (t1_ref.h:44-47).
GAIL_HD double rc_plane_scale(int nbps, int k) {
    const int plane = k == 0 ? nbps - 1 : nbps - 2 - (k - 1) / 3;
    return (double)(1ull << (2 * plane)) / 4096.0;
}

// Lower convex hull of a block's (rate, distortion) points (rate_control.cpp:38-86): len and
// dist are the block's cumulative truncation lengths and per-pass distortion reductions,
// scale is delta^2 * weight of its subband. Hands point j to out(j, point) and returns the
// number of points; their slopes fall strictly (the first may be infinite).
template <class Sink>
GAIL_HD int rc_block_hull_to(int np, int nbps, double scale, const uint32_t* len, const int32_t* dist, const Sink& out) {
    uint32_t R[kMaxPasses + 1];
    double D[kMaxPasses + 1];
    int K[kMaxPasses + 1];
    int m = 0;
    R[0] = 0;
    D[0] = 0.0;
    K[0] = 0;
    double acc = 0;
    for (int k = 0; k < np; ++k) {
        acc += (double)dist[k] * rc_plane_scale(nbps, k) * scale;
        const uint32_t r = len[k];
        if (acc <= D[m]) continue;
        while (m >= 1) {
            const double sNew = r == R[m] ? rc_inf() : (acc - D[m]) / (double)(r - R[m]);
            const double sLast = R[m] == R[m - 1] ? rc_inf() : (D[m] - D[m - 1]) / (double)(R[m] - R[m - 1]);
            if (sNew < sLast) break;
            --m;
        }
        ++m;
        R[m] = r;
        D[m] = acc;
        K[m] = k + 1;
    }
    for (int j = 1; j <= m; ++j) {
        const double dr = (double)R[j] - (double)R[j - 1];
        RcPoint q;
        q.slope = dr <= 0 ? rc_inf() : (D[j] - D[j - 1]) / dr;
        q.gain = D[j] - D[j - 1];
        q.rate = R[j];
        q.passes = (uint32_t)K[j];
        out(j - 1, q);
    }
    return m;
}

struct RcPointArray {
    RcPoint* p;
    GAIL_HD void operator()(int j, const RcPoint& q) const { p[j] = q; }
};

GAIL_HD int rc_block_hull(int np, int nbps, double scale, const uint32_t* len, const int32_t* dist, RcPoint* out) {
    return rc_block_hull_to(np, nbps, scale, len, dist, RcPointArray{out});
}

// Passes of a block kept at threshold lambda: up to its last hull point of slope >= lambda.
GAIL_HD uint32_t rc_select(const RcPoint* pts, int n, double lambda) {
    uint32_t np = 0;
    for (int h = 0; h < n; ++h) {
        if (pts[h].slope < lambda) break;
        np = pts[h].passes;
    }
    return np;
}

// ---------------------------------------------------------------- slope grid
// floor(1024 log2 s) of a finite s > 0 from its binary exponent and the 1,024 steps of one
// octave (mantissas of exp2(j / 1024); first[r]: largest j with mantissa[j] <= r << 42).
GAIL_HD int64_t rc_slope_bin(double s, const unsigned long long* mantissa, const uint16_t* first) {
    const unsigned long long bits = (unsigned long long)__builtin_bit_cast(unsigned long long, s);
    int e = (int)((bits >> 52) & 0x7ff);
    unsigned long long man = bits & ((1ull << 52) - 1);
    if (e == 0) {
        while (!(man & (1ull << 52))) {
            man <<= 1;
            --e;
        }
        man &= (1ull << 52) - 1;
        e += 1;
    }
    int j = first[man >> 42];
    while (j < 1023 && mantissa[j + 1] <= man) ++j;
    return (int64_t)(e - 1023) * 1024 + j;
}

GAIL_HD double rc_bin_edge(int64_t k, const unsigned long long* mantissa) {
    const int64_t oct = k >= 0 ? k / 1024 : -((-k + 1023) / 1024);
    const int j = (int)(k - oct * 1024);
    const unsigned long long bits = ((unsigned long long)(oct + 1023) << 52) | mantissa[j];
    return __builtin_bit_cast(double, bits);
}

// Bins of a frame's histogram: one octave beyond its slopes on either side, at most
// kRcMaxBins (the lowest bins give way).
GAIL_HD void rc_bin_range(double minSlope, double maxSlope, const unsigned long long* mantissa, const uint16_t* first,
                          int64_t& kLo, int& nb) {
    const int64_t kHi = rc_slope_bin(maxSlope, mantissa, first) + 1024;
    kLo = rc_slope_bin(minSlope, mantissa, first) - 1024;
    if (kHi - kLo + 2 > kRcMaxBins) kLo = kHi + 2 - kRcMaxBins;
    nb = (int)(kHi - kLo + 2);
}

GAIL_HD int rc_bin_of(double s, int64_t kLo, int nb, const unsigned long long* mantissa, const uint16_t* first) {
    if (!rc_finite(s)) return nb - 1;
    if (!(s > 0)) return 0;
    int64_t b = rc_slope_bin(s, mantissa, first) - kLo;
    return (int)(b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b));
}

// ---------------------------------------------------------------- packet headers
// B.10.1 packet-header bits: bytes fill MSB first and a byte after 0xFF carries 7 bits, so no
// marker can appear in a header. Each finished byte goes to `out` (RcCount drops it: the
// device only needs the count; the host's Tier-2 writer appends it to the codestream).
struct RcCount {
    GAIL_HD void operator()(uint32_t) const {}
};

template <class Out> struct RcBitsT {
    Out out;
    uint32_t acc, last, count;
    int used, room;
    GAIL_HD void init() {
        acc = last = count = 0;
        used = 0;
        room = 8;
    }
    GAIL_HD void emit() {
        last = acc & 0xFFu;
        out(last);
        ++count;
        room = last == 0xFFu ? 7 : 8;
        acc = 0;
        used = 0;
    }
    GAIL_HD void bits(uint32_t v, int n) {
        while (n > 0) {
            const int take = n < room - used ? n : room - used;
            n -= take;
            acc = (acc << take) | ((v >> n) & ((1u << take) - 1u));
            used += take;
            if (used == room) emit();
        }
    }
    // Pads the last byte with zeros; a header ending in 0xFF still owes its stuffed byte.
    GAIL_HD uint32_t finish() {
        if (used > 0 || count == 0) {
            acc <<= room - used;
            emit();
        }
        if (last == 0xFFu) {
            acc = 0;
            emit();
        }
        return count;
    }
};
using RcBits = RcBitsT<RcCount>;

GAIL_HD int rc_floor_log2(uint32_t v) { return 31 - __builtin_clz(v); }

// Table B.4: codewords for the number of coding passes.
template <class Sink> GAIL_HD void rc_pass_count(Sink& bw, int n) {
    if (n == 1) bw.bits(0, 1);
    else if (n == 2) bw.bits(2, 2);
    else if (n <= 5) bw.bits(0xCu | (uint32_t)(n - 3), 4);
    else if (n <= 36) bw.bits(0x1E0u | (uint32_t)(n - 6), 9);
    else bw.bits(0xFF80u | (uint32_t)(n - 37), 16);
}

// Codes one packet (one layer; B.10) with blk.passes(i) passes of frame-local block i: the
// header bits go to bw (after a leading 1: the packet is not empty), and body(i, len) is
// called for each included block in the order its bytes follow the header. Returns whether
// any block is included; an empty packet is instead the single byte 0x00, which the caller
// writes. blk gives passes(i), nbps(i) and len(i), the truncation length after its passes.
//
// The two tag trees of a precinct-band (B.10.2) are coded without node arrays. A DCI
// precinct-band holds at most 4 x 4 code-blocks, so a tree has at most three levels: the
// leaves, the 2 x 2 quads and the root. With one quality layer the coded bits follow from
// the node values and from which nodes were already visited (a bit mask per tree):
//   inclusion (threshold 1): walking root to leaf, a node not yet visited emits 1 if some
//     block below it is included (value 0) and 0 otherwise; the walk stops below a node of
//     value 1;
//   zero bit-planes (coded to the end, included blocks only): a node not yet visited emits
//     value - parent value zeros and a 1.
// That is what the general tag-tree coder emits for these thresholds. Everything else
// lives in registers (the leaves of a band are unrolled), so the device sizes a packet in
// one thread without memory traffic beyond its blocks.
template <class Blocks, class Sink, class Body>
GAIL_HD bool rc_packet_code(const RcPacket& P, const RcPBand* pbands, const Blocks& blk, Sink& bw, const Body& body) {
    bw.bits(1, 1);
    bool any = false;
    for (int k = 0; k < P.nband; ++k) {
        const RcPBand pb = pbands[P.first + k];
        const int w = pb.ncbx, h = pb.ncby;
        if (!w) continue;
        uint32_t np[16], len[16];
        int zv[16];
        uint32_t incl = 0;  // included leaves, bit y * 4 + x
GAIL_UNROLL
        for (int q = 0; q < 16; ++q) {
            np[q] = 0;
            len[q] = 0;
            zv[q] = 255;
            if ((q & 3) < w && (q >> 2) < h) {
                const uint32_t ci = pb.block0 + (uint32_t)(q >> 2) * pb.stride + (uint32_t)(q & 3);
                np[q] = blk.passes(ci);
                len[q] = blk.len(ci);
                zv[q] = (int)pb.mb - (int)blk.nbps(ci);
                if (np[q]) incl |= 1u << q;
            }
        }
        any = any || incl != 0;
        int zq[4];
GAIL_UNROLL
        for (int m = 0; m < 4; ++m) {
            const int q0 = (m >> 1) * 8 + (m & 1) * 2;
            const int a = zv[q0] < zv[q0 + 1] ? zv[q0] : zv[q0 + 1];
            const int b = zv[q0 + 4] < zv[q0 + 5] ? zv[q0 + 4] : zv[q0 + 5];
            zq[m] = a < b ? a : b;
        }
        const int zr0 = zq[0] < zq[1] ? zq[0] : zq[1], zr1 = zq[2] < zq[3] ? zq[2] : zq[3];
        const int zroot = zr0 < zr1 ? zr0 : zr1;
        const int levels = (w == 1 && h == 1) ? 1 : ((w <= 2 && h <= 2) ? 2 : 3);
        uint32_t ivis = 0, zvis = 0;  // visited nodes: leaves 0-15, quads 16-19, root 20
GAIL_UNROLL
        for (int q = 0; q < 16; ++q) {
            if ((q & 3) >= w || (q >> 2) >= h) continue;
            const int m = (q >> 3) * 2 + ((q >> 1) & 1);
            // inclusion tag tree
            bool go = true;
            if (levels >= 2) {
                const bool v0 = incl != 0;
                if (!(ivis & (1u << 20))) {
                    ivis |= 1u << 20;
                    bw.bits(v0 ? 1u : 0u, 1);
                }
                go = v0;
            }
            if (go && levels == 3) {
                const bool v0 = (incl & (0x33u << ((m >> 1) * 8 + (m & 1) * 2))) != 0;
                if (!(ivis & (1u << (16 + m)))) {
                    ivis |= 1u << (16 + m);
                    bw.bits(v0 ? 1u : 0u, 1);
                }
                go = v0;
            }
            if (go && !(ivis & (1u << q))) {
                ivis |= 1u << q;
                bw.bits(np[q] ? 1u : 0u, 1);
            }
            if (!np[q]) continue;
            // zero bit-planes tag tree
            int prev = 0;
            if (levels >= 2) {
                if (!(zvis & (1u << 20))) {
                    zvis |= 1u << 20;
                    bw.bits(1, zroot - prev + 1);
                }
                prev = zroot;
            }
            if (levels == 3) {
                if (!(zvis & (1u << (16 + m)))) {
                    zvis |= 1u << (16 + m);
                    bw.bits(1, zq[m] - prev + 1);
                }
                prev = zq[m];
            }
            if (!(zvis & (1u << q))) {
                zvis |= 1u << q;
                bw.bits(1, zv[q] - prev + 1);
            }
            rc_pass_count(bw, (int)np[q]);
            const int extra = rc_floor_log2(np[q]);
            const int need = len[q] ? rc_floor_log2(len[q]) + 1 : 1;
            const int incr = need - (3 + extra) > 0 ? need - (3 + extra) : 0;
            bw.bits(((1u << incr) - 1u) << 1, incr + 1);  // Lblock increment: incr ones, then 0
            bw.bits(len[q], 3 + incr + extra);
            body(pb.block0 + (uint32_t)(q >> 2) * pb.stride + (uint32_t)(q & 3), len[q]);
        }
    }
    return any;
}

struct RcBodySum {
    unsigned long long* bytes;
    GAIL_HD void operator()(uint32_t, uint32_t len) const { *bytes += len; }
};

// Bytes of one packet, header and body.
template <class Blocks>
GAIL_HD unsigned long long rc_packet_bytes(const RcPacket& P, const RcPBand* pbands, const Blocks& blk) {
    RcBits bw;
    bw.init();
    unsigned long long body = 0;
    if (!rc_packet_code(P, pbands, blk, bw, RcBodySum{&body})) return 1;
    return bw.finish() + body;
}

// ---------------------------------------------------------------- lambda search
// The search of rate_control.cpp:165-279 as a state machine. next() takes the sizes the
// last measurement gave (the first call: with every pass included) and returns true with
// lambda[] set when it wants another measurement, false when it is done. Histogram sums
// (bytes and included blocks per component, suffix sums over bins 0..nb) are read
// through data(c, b) / incl(c, b) of H.
struct RcSizes {
    unsigned long long total, comp[3];
};

struct RcSearch {
    // set by the caller
    int nb;
    int64_t kLo;
    double minSlope, maxSlope;
    unsigned long long maxTotal, maxComp;
    const unsigned long long* mantissa;
    const uint16_t* first;
    // state
    int stage, round, c;
    int bin[3];
    double lambda[3];
    double hdrCal[3], inclCal[3], beta[3];
    RcSizes sz;
    uint32_t iterations;

    GAIL_HD void init() {
        stage = 0;
        round = 0;
        c = 0;
        for (int k = 0; k < 3; ++k) {
            bin[k] = -1;
            lambda[k] = 0.0;
        }
        iterations = 0;
    }
    GAIL_HD double edge(int b) const { return rc_bin_edge(kLo + b, mantissa); }
    template <class H> GAIL_HD double hdata(const H& h, int cc, int b) const { return h.data(cc, b < 0 ? 0 : (b > nb ? nb : b)); }
    template <class H> GAIL_HD double hincl(const H& h, int cc, int b) const { return h.incl(cc, b < 0 ? 0 : (b > nb ? nb : b)); }
    template <class H> GAIL_HD void calibrate(const H& h) {
        for (int k = 0; k < 3; ++k) {
            const double data = hdata(h, k, bin[k]), incl = hincl(h, k, bin[k]);
            hdrCal[k] = (double)sz.comp[k] - data;
            inclCal[k] = incl;
            const double r = incl > 0 ? hdrCal[k] / incl : 3.0;
            beta[k] = incl > 0 ? (r < 1.0 ? 1.0 : (r > 8.0 ? 8.0 : r)) : 3.0;
        }
    }
    template <class H> GAIL_HD double model(const H& h, const int b[3], double comp[3]) const {
        double tot = (double)(sz.total - sz.comp[0] - sz.comp[1] - sz.comp[2]);
        for (int k = 0; k < 3; ++k) {
            comp[k] = hdata(h, k, b[k]) + hdrCal[k] + beta[k] * (hincl(h, k, b[k]) - inclCal[k]);
            tot += comp[k];
        }
        return tot;
    }
    // smallest bin in [lo, nb] whose model meets the budget (all components, or component cc)
    template <class H> GAIL_HD int search(const H& h, int cc, int lo) const {
        int hi = nb;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            int bb[3] = {bin[0], bin[1], bin[2]};
            double comp[3];
            bool ok;
            if (cc < 0) {
                bb[0] = bb[1] = bb[2] = mid;
                ok = model(h, bb, comp) <= (double)maxTotal;
            } else {
                bb[cc] = mid;
                model(h, bb, comp);
                ok = comp[cc] <= (double)maxComp;
            }
            if (ok) hi = mid;
            else lo = mid + 1;
        }
        return lo;
    }

    // Stages: 0 after the first measurement; 1 after a measurement of the main rounds;
    // 2 nudging the total; 3 choosing the next component over its cap; 4 after a measurement
    // of a component round; 5 nudging that component; 6 done.
    template <class H> GAIL_HD bool next(const H& h, const RcSizes& measured) {
        sz = measured;
        ++iterations;
        for (;;) {
            switch (stage) {
            case 0:
                if (sz.total > maxTotal) {
                    calibrate(h);
                    round = 0;
                    if (mainRound(h)) return true;
                    stage = 2;
                } else {
                    stage = 3;
                    c = 0;
                }
                break;
            case 1:
                if (sz.total <= maxTotal && sz.total + maxTotal / 500 >= maxTotal) {
                    stage = 2;
                    break;
                }
                calibrate(h);
                ++round;
                if (round < 6 && mainRound(h)) return true;
                stage = 2;
                break;
            case 2:
                if (sz.total > maxTotal && lambda[0] <= maxSlope) {
                    for (int k = 0; k < 3; ++k) lambda[k] = lambda[k] > 0 ? lambda[k] * 1.002 : minSlope;
                    return true;
                }
                stage = 3;
                c = 0;
                break;
            case 3:
                while (c < 3 && sz.comp[c] <= maxComp) ++c;
                if (c >= 3) {
                    stage = 6;
                    break;
                }
                for (int k = 0; k < 3; ++k) bin[k] = lambda[k] > 0 ? rc_bin_of(lambda[k], kLo, nb, mantissa, first) : -1;
                calibrate(h);
                round = 0;
                if (compRound(h)) return true;
                stage = 5;
                break;
            case 4:
                if (sz.comp[c] <= maxComp && sz.comp[c] + maxComp / 500 >= maxComp) {
                    stage = 5;
                    break;
                }
                calibrate(h);
                ++round;
                if (round < 6 && compRound(h)) return true;
                stage = 5;
                break;
            case 5:
                if (sz.comp[c] > maxComp && lambda[c] <= maxSlope) {
                    lambda[c] = lambda[c] > 0 ? lambda[c] * 1.002 : minSlope;
                    return true;
                }
                ++c;
                stage = 3;
                break;
            default:
                return false;
            }
        }
    }
    // one main round: false when the model has settled (the same bin again)
    template <class H> GAIL_HD bool mainRound(const H& h) {
        const int b = search(h, -1, 0);
        if (round > 0 && b == bin[0]) return false;
        for (int k = 0; k < 3; ++k) {
            bin[k] = b;
            lambda[k] = edge(b);
        }
        stage = 1;
        return true;
    }
    template <class H> GAIL_HD bool compRound(const H& h) {
        const int b = search(h, c, bin[c] > 0 ? bin[c] : 0);
        if (round > 0 && b == bin[c]) return false;
        bin[c] = b;
        lambda[c] = edge(b);
        stage = 4;
        return true;
    }
};

}  // namespace gail_amd
