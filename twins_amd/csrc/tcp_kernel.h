// Launch entry of the TCP slot kernels (tcp_kernel.hip), called by the batch runner.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tw
{
// a launch of the persistent world-queue kernel: `blocks` workgroups of `threads`, in teams of
// `team` workgroups per world; `hpw` hosts per wavefront work item; each world runs at most
// slot_budget slots per launch
struct LaunchShape
{
    int blocks, threads, team, hpw;
    int64_t slot_budget;
};
hipError_t tcp_launch(int runs, const void *batch_dev, const LaunchShape &shape, hipStream_t stream);
// workgroups of `threads` that one CU holds at once (the cooperative-launch limit)
hipError_t tcp_max_blocks(int runs, int threads, int *per_cu);
} // namespace tw
