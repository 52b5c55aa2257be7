// TEST INFRASTRUCTURE. The context-modelling kernel's whole-column pass formulation
// (libgail_amd/csrc/t1_cm_masks.h) run on the host, one lane per column, against a plain
// sequential EBCOT scan (ISO/IEC 15444-1 D.3: significance propagation, magnitude refinement,
// cleanup with run-length mode, every row visited in scan order). Both produce each pass's
// symbol bytes (context << 1 | bit); they must be equal for every pass of every block.
//
//   int t1cm_check(uint32_t seed, int nblocks, char* msg, int msglen, long long* stats, int plant)
//     random blocks (sizes 1..32, all four orientations, sparse to dense magnitudes);
//     returns the number of blocks whose symbols differ, msg describing the first. plant
//     puts a mistake into the whole-column side to show the comparison catches it: 1 the
//     below-left neighbour seen across a stripe boundary, 2 the fixed point stopped after
//     one round, 3 the signs above and below swapped.
//   int t1cm_zc_table_check(void)
//     zc_context against Table D.1 for every neighbourhood and orientation; returns the
//     number of disagreements.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "t1_cm_masks.h"
#include "t1_tables.h"

using namespace gail_amd;

namespace {

struct Block {
    int w, h, orient, nbps;
    uint32_t mag[32][32];  // [row][column]
    bool neg[32][32];
};

// ---------------------------------------------------------------- sequential scan
struct Seq {
    const Block& b;
    uint8_t zc[4 * 128];
    bool sig[32][32] = {}, visited[32][32] = {}, refined[32][32] = {};
    explicit Seq(const Block& blk) : b(blk) { build_zc_table(zc); }
    bool S(int y, int x) const { return y >= 0 && y < b.h && x >= 0 && x < b.w && sig[y][x]; }
    int chi(int y, int x) const { return S(y, x) ? (b.neg[y][x] ? -1 : 1) : 0; }
    int zcCtx(int y, int x) const {
        const int h = S(y, x - 1) + S(y, x + 1), v = S(y - 1, x) + S(y + 1, x);
        const int d = S(y - 1, x - 1) + S(y - 1, x + 1) + S(y + 1, x - 1) + S(y + 1, x + 1);
        return zc[b.orient * 128 + (h | v << 2 | d << 4)];
    }
    bool anyNb(int y, int x) const {
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx)
                if ((dy || dx) && S(y + dy, x + dx)) return true;
        return false;
    }
    uint8_t signSym(int y, int x) const {
        int hc = chi(y, x - 1) + chi(y, x + 1), vc = chi(y - 1, x) + chi(y + 1, x);
        hc = hc > 1 ? 1 : hc < -1 ? -1 : hc;
        vc = vc > 1 ? 1 : vc < -1 ? -1 : vc;
        int ctx, x0;
        if (hc == 0) {
            ctx = vc == 0 ? 9 : 10;
            x0 = vc < 0;
        } else {
            ctx = vc == 0 ? 12 : (vc == hc ? 13 : 11);
            x0 = hc < 0;
        }
        return (uint8_t)(ctx << 1 | ((b.neg[y][x] ? 1 : 0) ^ x0));
    }
    bool bit(int y, int x, int p) const { return (b.mag[y][x] >> p) & 1u; }
    void spp(int p, std::vector<uint8_t>& out) {
        for (int y0 = 0; y0 < b.h; y0 += 4)
            for (int x = 0; x < b.w; ++x)
                for (int y = y0; y < y0 + 4 && y < b.h; ++y) {
                    if (sig[y][x] || !anyNb(y, x)) continue;
                    const bool v = bit(y, x, p);
                    out.push_back((uint8_t)(zcCtx(y, x) << 1 | v));
                    visited[y][x] = true;
                    if (v) {
                        out.push_back(signSym(y, x));
                        sig[y][x] = true;
                    }
                }
    }
    void mr(int p, std::vector<uint8_t>& out) {
        for (int y0 = 0; y0 < b.h; y0 += 4)
            for (int x = 0; x < b.w; ++x)
                for (int y = y0; y < y0 + 4 && y < b.h; ++y) {
                    if (!sig[y][x] || visited[y][x]) continue;
                    const int ctx = refined[y][x] ? 16 : anyNb(y, x) ? 15 : 14;
                    out.push_back((uint8_t)(ctx << 1 | bit(y, x, p)));
                    refined[y][x] = true;
                }
    }
    void cleanup(int p, std::vector<uint8_t>& out) {
        for (int y0 = 0; y0 < b.h; y0 += 4)
            for (int x = 0; x < b.w; ++x) {
                int y = y0;
                if (y0 + 3 < b.h) {
                    bool run = true;
                    for (int k = 0; k < 4 && run; ++k)
                        run = !sig[y0 + k][x] && !visited[y0 + k][x] && !anyNb(y0 + k, x);
                    if (run) {
                        int first = 4;
                        for (int k = 0; k < 4; ++k)
                            if (bit(y0 + k, x, p)) {
                                first = k;
                                break;
                            }
                        out.push_back((uint8_t)(17 << 1 | (first < 4)));
                        if (first == 4) continue;
                        out.push_back((uint8_t)(18 << 1 | (first >> 1)));
                        out.push_back((uint8_t)(18 << 1 | (first & 1)));
                        out.push_back(signSym(y0 + first, x));
                        sig[y0 + first][x] = true;
                        y = y0 + first + 1;
                    }
                }
                for (; y < y0 + 4 && y < b.h; ++y) {
                    if (sig[y][x] || visited[y][x]) continue;
                    const bool v = bit(y, x, p);
                    out.push_back((uint8_t)(zcCtx(y, x) << 1 | v));
                    if (v) {
                        out.push_back(signSym(y, x));
                        sig[y][x] = true;
                    }
                }
            }
        std::memset(visited, 0, sizeof visited);
    }
};

// ---------------------------------------------------------------- whole-column passes
// The kernel's formulation: lanes are columns; from_left / from_right read the neighbouring
// lane (zero at the block's edge); symbols of a pass are counted per stripe, placed by an
// exclusive prefix over the columns, stripe after stripe.
struct Cols {
    const Block& b;
    int plant = 0;
    uint32_t S[32] = {}, R[32] = {}, V[32] = {}, sgn[32] = {}, colMask[32] = {}, sig0[32] = {};
    explicit Cols(const Block& blk) : b(blk) {
        const uint32_t rows = b.h >= 32 ? 0xFFFFFFFFu : ((1u << b.h) - 1u);
        for (int x = 0; x < 32; ++x) {
            colMask[x] = x < b.w ? rows : 0u;
            for (int y = 0; y < b.h && x < b.w; ++y) sgn[x] |= (b.neg[y][x] ? 1u : 0u) << y;
        }
    }
    static uint32_t fromLeft(const uint32_t* v, int x) { return x == 0 ? 0u : v[x - 1]; }
    static uint32_t fromRight(const uint32_t* v, int x) { return x == 31 ? 0u : v[x + 1]; }
    uint32_t plane(int x, int p) const {
        uint32_t m = 0;
        for (int y = 0; y < b.h && x < b.w; ++y) m |= ((b.mag[y][x] >> p) & 1u) << y;
        return m;
    }
    static uint32_t bytes(uint32_t m, int r0) { return (((m >> r0) & 0xFu) * 0x204081u) & 0x01010101u; }

    // type 0 SPP, 1 MR, 2 cleanup
    void pass(int type, int p, std::vector<uint8_t>& out, int* rounds) {
        uint32_t Bp[32], keep0[32] = {}, keep1[32] = {}, runs[32] = {}, rl[32] = {};
        uint32_t w0m[32][5] = {}, w1m[32][4] = {}, add0 = 0;
        for (int x = 0; x < 32; ++x) Bp[x] = plane(x, p);
        if (type == 0) std::memcpy(sig0, S, sizeof S);  // significant before this plane
        uint32_t SL[32], SR[32];
        for (int x = 0; x < 32; ++x) {
            SL[x] = fromLeft(S, x);
            SR[x] = fromRight(S, x);
        }
        if (type == 1) {
            for (int x = 0; x < 32; ++x) {
                const uint32_t nAny = neighbour_any(S[x], SL[x], SR[x]);
                keep0[x] = sig0[x] & colMask[x];
                w0m[x][0] = Bp[x];
                w0m[x][1] = nAny & ~R[x];
                w0m[x][2] = R[x];
            }
            add0 = 0x1C1C1C1Cu;
        } else {
            uint32_t N[32] = {}, E[32] = {}, open[32] = {}, around[32] = {};
            if (type == 0) {
                for (int x = 0; x < 32; ++x) {
                    open[x] = ~S[x] & colMask[x];
                    around[x] = neighbour_any(S[x], SL[x], SR[x]);
                }
                // least fixed point over the columns
                int r = 0;
                for (;;) {
                    ++r;
                    if (plant == 2 && r > 1) break;
                    uint32_t turned[32];
                    bool change = false;
                    for (int x = 0; x < 32; ++x) {
                        const uint32_t seen = around[x] | new_seen(fromLeft(N, x), fromRight(N, x));
                        E[x] = column_fill(open[x] & seen, open[x], Bp[x]);
                        turned[x] = E[x] & Bp[x];
                        change |= turned[x] != N[x];
                    }
                    if (!change) break;
                    std::memcpy(N, turned, sizeof N);
                }
                if (rounds) *rounds += r;
                for (int x = 0; x < 32; ++x) {
                    keep0[x] = E[x];
                    keep1[x] = N[x];
                }
            } else {
                for (int x = 0; x < 32; ++x) {
                    const uint32_t Cc = colMask[x] & ~S[x] & ~V[x];
                    N[x] = Cc & Bp[x];
                }
                for (int x = 0; x < 32; ++x) {
                    const uint32_t Cc = colMask[x] & ~S[x] & ~V[x];
                    runs[x] = run_stripes(Cc, S[x], SL[x], SR[x], N[x], fromLeft(N, x), fromRight(N, x));
                    cleanup_run_symbols(Cc, N[x], runs[x], keep0[x], rl[x]);
                    keep1[x] = N[x];
                }
            }
            for (int x = 0; x < 32; ++x) {
                Seen v = seen_at_coding(S[x], SL[x], SR[x], N[x], fromLeft(N, x), fromRight(N, x));
                if (plant == 1) v.DL = (SL[x] | fromLeft(N, x)) >> 1;
                const ZcCtx z = zc_context(v, b.orient == 1, b.orient == 3);
                const uint32_t sU = plant == 3 ? sgn[x] >> 1 : sgn[x] << 1, sD = plant == 3 ? sgn[x] << 1 : sgn[x] >> 1;
                const ScCtx s = sign_context(v, fromLeft(sgn, x), fromRight(sgn, x), sgn[x], sU, sD);
                w0m[x][0] = Bp[x];
                w0m[x][1] = z.c0;
                w0m[x][2] = z.c1;
                w0m[x][3] = z.c2;
                w0m[x][4] = z.c3;
                w1m[x][0] = s.d;
                w1m[x][1] = s.c0;
                w1m[x][2] = s.c1;
                w1m[x][3] = s.c2;
            }
            for (int x = 0; x < 32; ++x) {
                S[x] |= N[x];
                if (type == 0) V[x] |= E[x];
            }
        }
        if (type == 1)
            for (int x = 0; x < 32; ++x) R[x] |= keep0[x];
        // counts per stripe, exclusive prefix over the columns, stripe after stripe
        uint32_t cnt[32];
        for (int x = 0; x < 32; ++x) cnt[x] = nibble_counts(keep0[x]) + nibble_counts(keep1[x]) + rl[x];
        const size_t base = out.size();
        uint32_t stripeBase = 0;
        for (int s = 0; s < 8; ++s) {
            const int r0 = 4 * s;
            uint32_t excl = 0;
            for (int x = 0; x < 32; ++x) {
                const uint32_t n = (cnt[x] >> r0) & 0xFu;
                uint32_t pos = (uint32_t)base + stripeBase + excl;
                excl += n;
                if (n == 0) continue;
                if (out.size() < pos + n) out.resize(pos + n, 0xEE);
                const uint32_t z = (keep0[x] >> r0) & 0xFu, g = (keep1[x] >> r0) & 0xFu;
                uint32_t w0 = add0, w1 = 0x12121212u;
                for (int i = 0; i < 5; ++i) w0 += bytes(w0m[x][i], r0) << i;
                for (int i = 0; i < 4; ++i) w1 += bytes(w1m[x][i], r0) << i;
                if ((runs[x] >> r0) & 1u) {
                    out[pos++] = (uint8_t)(34u | (g != 0));
                    if (g) {
                        const uint32_t first = (uint32_t)__builtin_ctz(g);
                        out[pos++] = (uint8_t)(36u | (first >> 1));
                        out[pos++] = (uint8_t)(36u | (first & 1u));
                    }
                }
                for (int k = 0; k < 4; ++k) {
                    if ((z >> k) & 1u) out[pos++] = (uint8_t)(w0 >> (8 * k));
                    if ((g >> k) & 1u) out[pos++] = (uint8_t)(w1 >> (8 * k));
                }
            }
            stripeBase += excl;
        }
        if (type == 2) std::memset(V, 0, sizeof V);
    }
};

Block randomBlock(std::mt19937& rng) {
    Block b{};
    std::uniform_int_distribution<int> size(1, 32), kind(0, 5), orient(0, 3);
    const int k = kind(rng);
    b.w = k == 0 ? 32 : size(rng);
    b.h = k == 1 ? 32 : size(rng);
    if (k == 2) b.w = b.h = 32;
    b.orient = orient(rng);
    // magnitude distributions: sparse large values, dense small values, Laplacian-like,
    // smooth ramps with noise, isolated points, everything
    const int dist = (int)(rng() % 6);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    int maxbits = 1;
    for (int y = 0; y < b.h; ++y)
        for (int x = 0; x < b.w; ++x) {
            uint32_t m = 0;
            switch (dist) {
            case 0: m = u(rng) < 0.05 ? (uint32_t)(rng() % 4096) : 0; break;
            case 1: m = (uint32_t)(rng() % 4); break;
            case 2: m = (uint32_t)(-std::log(1.0 - u(rng) * 0.999999) * 20.0); break;
            case 3: m = (uint32_t)((x * 7 + y * 3) % 50 + rng() % 5); break;
            case 4: m = ((x * 131 + y * 17) % 97 == 0) ? 1000u + (uint32_t)(rng() % 1000) : 0; break;
            default: m = (uint32_t)(rng() % 65536); break;
            }
            b.mag[y][x] = m;
            b.neg[y][x] = m && (rng() & 1);
            while ((m >> maxbits) != 0) ++maxbits;
        }
    b.nbps = maxbits;
    return b;
}

}  // namespace

extern "C" int t1cm_zc_table_check(void) {
    uint8_t zc[4 * 128];
    build_zc_table(zc);
    int bad = 0;
    for (int o = 0; o < 4; ++o)
        for (int nb = 0; nb < 256; ++nb) {
            // one row (bit 1) and its eight neighbours as whole-column masks
            Seen v{};
            const uint32_t r = 1u << 1;
            v.L = (nb & 1) ? r : 0;
            v.R = (nb & 2) ? r : 0;
            v.U = (nb & 4) ? r : 0;
            v.D = (nb & 8) ? r : 0;
            v.UL = (nb & 16) ? r : 0;
            v.DL = (nb & 32) ? r : 0;
            v.UR = (nb & 64) ? r : 0;
            v.DR = (nb & 128) ? r : 0;
            const ZcCtx c = zc_context(v, o == 1, o == 3);
            const int got = (int)(((c.c0 >> 1) & 1) | ((c.c1 >> 1) & 1) << 1 | ((c.c2 >> 1) & 1) << 2 | ((c.c3 >> 1) & 1) << 3);
            const int h = (nb & 1) + ((nb >> 1) & 1), vv = ((nb >> 2) & 1) + ((nb >> 3) & 1);
            const int d = ((nb >> 4) & 1) + ((nb >> 5) & 1) + ((nb >> 6) & 1) + ((nb >> 7) & 1);
            const int want = zc[o * 128 + (h | vv << 2 | d << 4)];
            bad += got != want;
        }
    return bad;
}

extern "C" int t1cm_check(uint32_t seed, int nblocks, char* msg, int msglen, long long* stats, int plant) {
    std::mt19937 rng(seed);
    int bad = 0;
    long long symbols = 0, passes = 0, rounds = 0;
    for (int i = 0; i < nblocks; ++i) {
        const Block b = randomBlock(rng);
        Seq seq(b);
        Cols cols(b);
        cols.plant = plant;
        for (int p = b.nbps - 1; p >= 0; --p) {
            for (int type = (p == b.nbps - 1 ? 2 : 0); type < 3; ++type) {
                std::vector<uint8_t> a, c;
                if (type == 0) seq.spp(p, a);
                if (type == 1) seq.mr(p, a);
                if (type == 2) seq.cleanup(p, a);
                int r = 0;
                cols.pass(type, p, c, &r);
                rounds += r;
                ++passes;
                symbols += (long long)a.size();
                if (a != c) {
                    if (!bad && msg) {
                        size_t k = 0;
                        while (k < a.size() && k < c.size() && a[k] == c[k]) ++k;
                        std::snprintf(msg, (size_t)msglen,
                                      "block %d (%dx%d orient %d nbps %d) plane %d pass %d: %zu vs %zu symbols, first "
                                      "difference at %zu (%d vs %d)",
                                      i, b.w, b.h, b.orient, b.nbps, p, type, a.size(), c.size(), k,
                                      k < a.size() ? a[k] : -1, k < c.size() ? c[k] : -1);
                    }
                    ++bad;
                    p = -1;
                    break;
                }
            }
        }
    }
    if (stats) {
        stats[0] = symbols;
        stats[1] = passes;
        stats[2] = rounds;
    }
    return bad;
}
