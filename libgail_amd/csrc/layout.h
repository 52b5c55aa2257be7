// Frame geometry of a single-tile DCI JPEG 2000 frame: subbands in Mallat layout, 32x32
// code-blocks with ragged edges and the quantisation step of every subband (the values of
// the reference's layout.cpp:77-210; precincts are the packet tables' business).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gail_amd {

constexpr int kComps = 3;
constexpr int kCblkLog2 = 5;
constexpr int kCblk = 32;
constexpr int kBitDepth = 12;

// Same member order and types as the reference's gail::Subband (it crosses the plugin
// boundary as plain data).
struct Subband {
    int comp = 0, res = 0, level = 0, orient = 0;  // orient: 0 LL, 1 HL, 2 LH, 3 HH
    int px0 = 0, py0 = 0;
    int w = 0, h = 0;
    int ncbx = 0, ncby = 0;
    int cblkBase = 0;
    int eps = 0, mu = 0;
    int Mb = 0;
    float delta = 1.f;
    double weight = 1.0;
};

struct CodeBlock {
    int band = 0;
    int x0 = 0, y0 = 0, w = 0, h = 0;
};

struct Layout {
    bool fourk = false;
    int width = 0, height = 0, levels = 5, guardBits = 1;
    double baseStep = 1.0;
    std::vector<Subband> bands;
    std::vector<CodeBlock> cblks;

    int numRes() const { return levels + 1; }
    int bandsPerComp() const { return 3 * levels + 1; }
    int numCblks() const { return (int)cblks.size(); }
    int bandIndex(int comp, int res, int orient) const {
        return comp * bandsPerComp() + (res == 0 ? 0 : 1 + 3 * (res - 1) + (orient - 1));
    }
};

// Geometry only (no step sizes). Bands are ordered component by component, LL first, then
// HL, LH, HH of each resolution from the coarsest; blocks band by band in raster order.
Layout makeGeometry(int width, int height, bool fourk, int levels, int guardBits);
// Full layout with the profile's levels / guard bits and steps chosen from baseStep.
// Throws std::runtime_error on an invalid size.
Layout makeLayout(int width, int height, bool fourk, double baseStep);
// Applies signalled (eps, mu) pairs (one per band of a component) to every component.
void setSignaledSteps(Layout& L, const int* eps, const int* mu);
double synthesisNorm(int level, int orient);

// DCI container sizes: flat 1998x1080 / 3996x2160, scope 2048x858 / 4096x1716,
// full 2048x1080 / 4096x2160. container: 0 flat, 1 scope, 2 full.
void containerSize(bool fourk, int container, int& w, int& h);

}  // namespace gail_amd
