// ff_fused.hip -- fused per-level kernels (in development).
#include <hip/hip_runtime.h>

#include "ff_device.hpp"
#include "ff_kernels.hpp"

namespace ffamd {

bool fused_level_supported(const LevelArgs&) { return false; }

void launch_fused_level(const LevelArgs&, hipStream_t) {}

}  // namespace ffamd
