// Synthetic DCI test frames: three planar uint16 X'Y'Z' channels holding 12-bit codes.
// Each channel, from its own seed: 4 summed 2-D cosines (frequency 0.5-40 cycles per
// frame in a random direction, amplitude 100-600, random phase), 40 axis-aligned
// rectangles with uniform random levels in [-512, 512], Gaussian noise of sigma 30, then
// offset to a mean of 2048, rounded and clipped to [0, 4095].
//
// Deterministic for a (seed, width, height) on any host: a counter-based generator
// (splitmix64 of the seed and the sample index) drives the noise, so rows can be filled
// on any number of threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "gail_amd_export.h"

namespace {

uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Stream {
    uint64_t s;
    double uniform() { return (double)(splitmix(s++) >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
};

// Standard normal from two uniforms of the counter stream (Box-Muller).
double gauss(uint64_t key) {
    const double u1 = ((double)(splitmix(key) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(splitmix(key ^ 0xD1B54A32D192ED03ull) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

void channel(uint64_t seed, int W, int H, uint16_t* out, int threads) {
    Stream rs{splitmix(seed)};
    struct Cos {
        double fx, fy, amp, phase;
    } cs[4];
    for (auto& c : cs) {
        const double f = rs.range(0.5, 40.0), dir = rs.range(0.0, 6.283185307179586);
        c.fx = f * std::cos(dir) / W;
        c.fy = f * std::sin(dir) / H;
        c.amp = rs.range(100.0, 600.0);
        c.phase = rs.range(0.0, 6.283185307179586);
    }
    struct Rect {
        int x0, y0, x1, y1;
        double level;
    } rects[40];
    for (auto& r : rects) {
        int a = (int)(rs.uniform() * W), b = (int)(rs.uniform() * W);
        int c = (int)(rs.uniform() * H), d = (int)(rs.uniform() * H);
        r = {std::min(a, b), std::min(c, d), std::max(a, b) + 1, std::max(c, d) + 1, rs.range(-512.0, 512.0)};
    }
    // Separable cosines: cos(a + b) = cos a cos b - sin a sin b.
    std::vector<double> cx(4 * (size_t)W), sx(4 * (size_t)W), cy(4 * (size_t)H), sy(4 * (size_t)H);
    for (int k = 0; k < 4; ++k) {
        for (int x = 0; x < W; ++x) {
            const double a = 6.283185307179586 * cs[k].fx * x + cs[k].phase;
            cx[k * (size_t)W + x] = cs[k].amp * std::cos(a);
            sx[k * (size_t)W + x] = cs[k].amp * std::sin(a);
        }
        for (int y = 0; y < H; ++y) {
            const double b = 6.283185307179586 * cs[k].fy * y;
            cy[k * (size_t)H + y] = std::cos(b);
            sy[k * (size_t)H + y] = std::sin(b);
        }
    }
    std::vector<double> field((size_t)W * H);
    const uint64_t noiseKey = splitmix(seed ^ 0x6A09E667F3BCC909ull);
    std::atomic<int> next{0};
    auto rows = [&] {
        for (int y; (y = next.fetch_add(1)) < H;) {
            double* f = &field[(size_t)y * W];
            for (int x = 0; x < W; ++x) {
                double v = 0;
                for (int k = 0; k < 4; ++k)
                    v += cx[k * (size_t)W + x] * cy[k * (size_t)H + y] - sx[k * (size_t)W + x] * sy[k * (size_t)H + y];
                f[x] = v + 30.0 * gauss(noiseKey + (uint64_t)y * W + x);
            }
            for (const Rect& r : rects)
                if (y >= r.y0 && y < r.y1)
                    for (int x = r.x0; x < std::min(r.x1, W); ++x) f[x] += r.level;
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < threads; ++t) ts.emplace_back(rows);
    rows();
    for (auto& t : ts) t.join();
    double sum = 0;
    for (double v : field) sum += v;
    const double off = 2048.0 - sum / (double)field.size();
    for (size_t i = 0; i < field.size(); ++i) {
        const double v = std::floor(field[i] + off + 0.5);
        out[i] = (uint16_t)std::min(4095.0, std::max(0.0, v));
    }
}

}  // namespace

// Fills three planes of W x H codes (plane c at out + c * W * H).
extern "C" GAIL_AMD_EXPORT void gail_synth_xyz12(uint64_t seed, int width, int height, uint16_t* out, int threads) {
    for (int c = 0; c < 3; ++c)
        channel(splitmix(seed * 3 + (uint64_t)c + 1), width, height, out + (size_t)c * width * height,
                std::max(1, threads));
}
