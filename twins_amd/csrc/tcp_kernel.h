// Launch entry of the TCP slot kernels (tcp_kernel.hip), called by the batch runner.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tw
{
// one launch of the persistent world-queue kernel over a packed batch (BatchT<runs> in device
// memory); each world runs at most slot_budget slots in this launch
hipError_t tcp_launch(int runs, const void *batch_dev, int blocks, int threads, int64_t slot_budget, hipStream_t stream);
} // namespace tw
