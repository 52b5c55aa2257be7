// The reference library's side of the HIP backend: gail::Device, gail::EncoderSlot and
// gail::DecoderSlot (src/core/backend.h:47-93, src/core/decode_backend.h:42-51) over the
// plugin table of include/gail_hip_plugin.h, the counterpart of the CUDA host's plugin
// wrappers (src/cuda/cuda_host.cpp:152-258) and loader (:82-93, :386-420). It is written
// against the reference's headers and meant to sit in its tree as src/hip/hip_host.{h,cpp};
// here it is compiled against them where they are (oracle/Makefile).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "core/backend.h"

namespace gail {

struct HipPluginInfo {
    std::string path;     // the plugin opened, or tried
    std::string problem;  // "" when it loaded and passed the interface checks
    std::string toolkit;  // the HIP runtime it was built with
};

// Loads libgail_hip.so from `dir` (the gail library's own directory) once per process and
// checks its ABI version, table size and structure layout hash.
HipPluginInfo loadHipPlugin(const std::string& dir);

// The HIP devices the plugin can use (one GPU in this build: index 0 when there is one),
// set up and ready to make slots. Devices it cannot set up are left out with a note to log.
std::vector<std::unique_ptr<Device>> createHipDevices(const std::string& dir, const std::vector<int>& select,
                                                      const Logger& log);

}  // namespace gail
