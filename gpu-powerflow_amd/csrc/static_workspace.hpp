// Fixed device workspace of the allocation-free inner-solver entry point.
#pragma once

#include <cstddef>

namespace cupf {

// Size of the per-device workspace (128 MiB).
constexpr size_t kStaticWorkspaceBytes = size_t(128) << 20;

// Device address of the workspace on the current device, looked up on the
// first call per device and cached.  Every call on one device returns the
// same memory: work that overlaps on one GPU races.
void* static_workspace();

}  // namespace cupf
