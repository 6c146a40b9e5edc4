// Launch entry of the TCP slot kernels (tcp_kernel.hip), called by the batch runner.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tw
{
// a launch of the persistent world-queue kernel: `blocks` workgroups of `threads` (their teams,
// sizes and queues are the batch's TeamPlan); `teams` when a team has several workgroups (the
// grid must then fit the GPU at once); each world runs at most slot_budget slots per launch
struct LaunchShape
{
    int blocks, threads, teams, pad;
    int64_t slot_budget;
};
hipError_t tcp_launch(int runs, const void *batch_dev, const LaunchShape &shape, hipStream_t stream);
// the serving loop's persistent kernel over a ServerCtl in device memory
hipError_t tcp_server_launch(int runs, const void *ctl_dev, int blocks, int threads, hipStream_t stream);
// zeroes the world-queue counters and the team barrier state before a launch
hipError_t tcp_reset(uint32_t *queue_next, int nq, void *teams, size_t team_bytes, hipStream_t stream);
// workgroups of `threads` that one CU holds at once (times the CUs: the limit of a plan with teams)
hipError_t tcp_max_blocks(int runs, int threads, int *per_cu);
} // namespace tw
