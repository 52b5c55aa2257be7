// What libgail.so's encoder and server share: the backend plugin, loaded from this
// library's directory, and the per-stream set-up taken from a gail_config (frame layout,
// the plugin's layout and settings, rate limits, and the copy of an app frame into the
// plugin's input layout, gail_api.cpp:363-396, 566-739).
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "codestream.h"
#include "gail_amd.h"
#include "gail_hip_plugin.h"
#include "layout.h"

namespace gail_amd {

// The thread's last error message (gail_last_error).
std::string& lastErrorText();

struct ApiError : std::runtime_error {
    int status;
    ApiError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

struct PluginHandle {
    const GailHipPlugin* table = nullptr;
    std::string error;
};
// libgail_hip.so next to this library, ABI / size / layout checked (cuda_host.cpp:82-93, 386-420).
const PluginHandle& loadPlugin();

bool allZero(const uint64_t* p, int n);

// Device memory left to other users of the GPU: max(512 MiB, 6 % of it) (pipeline.cpp:19).
uint64_t memoryHeadroom(uint64_t total);
// Frames in flight that fit: as many as requested (at least one), fewer when `available`
// less the headroom, or `limit` (0: none), holds fewer of perFrame bytes plus a tenth of
// slack; available 0 means unknown (pipeline.cpp:21-34).
int fitFrames(uint64_t available, uint64_t total, uint64_t perFrame, int requested, uint64_t limit);

struct PlanarInfo {
    int bits, sx, sy, rgb;
};
bool planarInfo(gail_pixel_format f, PlanarInfo& p);

// One stream's geometry, input and limits. Never copied or moved once made (plainLayout
// points into bands).
struct StreamSpec {
    gail_config cfg{};
    Layout L;
    std::vector<GailHipSubband> bands;
    GailHipLayout plainLayout{};
    GailHipEncodeSettings settings{};
    RateLimits limits;
    int width = 0, height = 0, picW = 0, picH = 0, ox = 0, oy = 0;
    size_t inBytes = 0;
    // planar inputs: where each plane lies in the plugin's input layout (planes back to back,
    // rows unpadded) and its rows
    struct Plane {
        size_t offset = 0, rowBytes = 0;
        int rows = 0;
    };
    bool planar = false;
    Plane planes[3];
    double slopeFactor = 0.25;  // gail_api.cpp:607
    std::vector<float> lut3d;   // the stream's copy of an app 3D LUT (settings point here)

    // Validates c and builds the stream (throws ApiError).
    static std::unique_ptr<StreamSpec> make(const gail_config& c);
    StreamSpec(const StreamSpec&) = delete;
    StreamSpec& operator=(const StreamSpec&) = delete;
    // Checks that a frame has the planes this stream reads.
    int checkFrame(const gail_frame& f, std::string& why) const;
    // Copies an app frame into the plugin's input layout (inBytes at dst).
    void fill(const gail_frame& f, uint8_t* dst) const { fill(f, dst, 0, 1); }
    // Part `part` of `parts` of that copy: the rows [rows * part / parts, rows * (part + 1) /
    // parts) of every plane. Parts of one frame may run at once on different threads when the
    // picture fills the container (splitParts()); otherwise only fill(f, dst, 0, 1).
    void fill(const gail_frame& f, uint8_t* dst, int part, int parts) const;
    bool splitParts() const { return picW == width && picH == height; }
    // When the frame already is the plugin's input layout in one block of memory, that
    // block (inBytes long); else nullptr.
    const void* directInput(const gail_frame& f) const;
    bool fourk() const { return cfg.profile == GAIL_PROFILE_4K; }
    double eyeFps() const { return cfg.fps * (cfg.stereo ? 2 : 1); }

private:
    StreamSpec() = default;
};

}  // namespace gail_amd
