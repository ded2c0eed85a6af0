// device_build.h -- BLAS tree construction on the device.
#ifndef RDRT_DEVICE_BUILD_H
#define RDRT_DEVICE_BUILD_H

#include <hip/hip_runtime.h>

#include <memory>

#include "scene.h"

namespace rdrt {

// A DeviceTreeBuilder on `stream` of the current device.
std::unique_ptr<DeviceTreeBuilder> make_device_tree_builder(hipStream_t stream);

}  // namespace rdrt

#endif
