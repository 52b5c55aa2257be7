// C entry points of libgail.so for frame geometry: what the library computes before it
// creates an encoder (the reference's makeLayout, layout.cpp:181-210, and
// containerSize, layout.cpp:221-228), in the plain form the plugin boundary takes.
#include <algorithm>
#include <cstring>
#include <vector>

#include "color_tables.h"
#include "gail_amd_export.h"
#include "gail_hip_plugin.h"
#include "layout.h"
#include "stream.h"

using namespace gail_amd;

// Fills up to `cap` subbands; returns the number of code-blocks, or -1 for an invalid size.
extern "C" GAIL_AMD_EXPORT int gail_layout_make(int width, int height, int fourk, double baseStep,
                                                 GailHipSubband* bands, int cap, int* nbands) {
    try {
        const Layout L = makeLayout(width, height, fourk != 0, baseStep);
        if (nbands) *nbands = (int)L.bands.size();
        for (int i = 0; i < (int)L.bands.size() && i < cap; ++i) {
            const Subband& b = L.bands[i];
            GailHipSubband& o = bands[i];
            std::memset(&o, 0, sizeof o);
            o.comp = b.comp;
            o.res = b.res;
            o.level = b.level;
            o.orient = b.orient;
            o.px0 = b.px0;
            o.py0 = b.py0;
            o.w = b.w;
            o.h = b.h;
            o.ncbx = b.ncbx;
            o.ncby = b.ncby;
            o.cblkBase = b.cblkBase;
            o.eps = b.eps;
            o.mu = b.mu;
            o.Mb = b.Mb;
            o.delta = b.delta;
            o.weight = b.weight;
        }
        return L.numCblks();
    } catch (...) {
        return -1;
    }
}

extern "C" GAIL_AMD_EXPORT int gail_layout_cblks(int width, int height, int fourk, int* out5, int cap) {
    try {
        const Layout L = makeLayout(width, height, fourk != 0, 1.0);
        for (int i = 0; i < L.numCblks() && i < cap; ++i) {
            const CodeBlock& c = L.cblks[i];
            int* o = out5 + 5 * i;
            o[0] = c.band;
            o[1] = c.x0;
            o[2] = c.y0;
            o[3] = c.w;
            o[4] = c.h;
        }
        return L.numCblks();
    } catch (...) {
        return -1;
    }
}

// The .cube reader the plugin uses for an app LUT file: fills up to cap floats at out and
// the LUT's size; returns the float count, or -1 with the reason in msg.
extern "C" GAIL_AMD_EXPORT int64_t gail_read_cube_lut(const char* path, float* out, uint64_t cap, int* size, char* msg,
                                                      int msgCap) {
    try {
        int n = 0;
        const std::vector<float> v = readCubeLut(path, n);
        *size = n;
        std::memcpy(out, v.data(), sizeof(float) * std::min<uint64_t>(cap, v.size()));
        return (int64_t)v.size();
    } catch (const std::exception& e) {
        if (msg && msgCap > 0) {
            std::strncpy(msg, e.what(), (size_t)msgCap - 1);
            msg[msgCap - 1] = 0;
        }
        return -1;
    }
}

// The frames-in-flight rule shared by the encoder and the server (memoryHeadroom, fitFrames).
extern "C" GAIL_AMD_EXPORT uint64_t gail_memory_headroom(uint64_t total) { return memoryHeadroom(total); }
extern "C" GAIL_AMD_EXPORT int gail_fit_frames(uint64_t available, uint64_t total, uint64_t perFrame, int requested,
                                               uint64_t limit) {
    return fitFrames(available, total, perFrame, requested, limit);
}
