// Error kinds that cross the plugin boundary as distinct status codes.
#pragma once
#include <stdexcept>

namespace gail_amd {

// Device (or pinned host) memory for a slot or job could not be allocated; the device
// stays usable and the slots created so far keep working (backend.h:68-72).
struct DeviceOutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace gail_amd
