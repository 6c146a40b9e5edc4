// Launch entry of the TCP slot kernels (tcp_kernel.hip), called by the batch runner.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace tw
{
// the longest a serving-loop team holds a claim on a request not yet published (seconds); an idle
// team gives its claim back sooner (the server's idle_s, which must be less)
constexpr double SERVER_WAIT_S = 600.0;
// a launch of the persistent world-queue kernel: `blocks` workgroups of `threads` (their teams,
// sizes and queues are the batch's TeamPlan); `teams` when a team has several workgroups (the
// grid must then fit the GPU at once); each world runs at most slot_budget slots per launch
struct LaunchShape
{
    int blocks, threads, teams, pad;
    int64_t slot_budget;
};
// per chain length (tcp_kernel.hip is built once for 9 runs, once for 25)
hipError_t tcp_launch_9(const void *batch_dev, const LaunchShape &shape, hipStream_t stream);
hipError_t tcp_launch_25(const void *batch_dev, const LaunchShape &shape, hipStream_t stream);
hipError_t tcp_server_launch_9(const void *ctl_dev, int blocks, int threads, hipStream_t stream);
hipError_t tcp_server_launch_25(const void *ctl_dev, int blocks, int threads, hipStream_t stream);
hipError_t tcp_max_blocks_9(int threads, int *per_cu);
hipError_t tcp_max_blocks_25(int threads, int *per_cu);
// zeroes the world-queue counters and the team barrier state before a launch
hipError_t tcp_reset(uint32_t *queue_next, int nq, void *teams, size_t team_bytes, hipStream_t stream);
// the library's static scratch array on the current device (tw_batch_run_noalloc): *ptr is its
// device address, looked up once per device
constexpr size_t TW_SCRATCH_BYTES = 64ull << 20;
hipError_t tcp_static_scratch(void **ptr);

// the persistent world-queue kernel over a batch (BatchT<runs> in device memory)
inline hipError_t tcp_launch(int runs, const void *batch_dev, const LaunchShape &shape, hipStream_t stream)
{
    return runs == 25 ? tcp_launch_25(batch_dev, shape, stream) : tcp_launch_9(batch_dev, shape, stream);
}
// the serving loop's persistent kernel over a ServerCtl in device memory
inline hipError_t tcp_server_launch(int runs, const void *ctl_dev, int blocks, int threads, hipStream_t stream)
{
    return runs == 25 ? tcp_server_launch_25(ctl_dev, blocks, threads, stream) : tcp_server_launch_9(ctl_dev, blocks, threads, stream);
}
// workgroups of `threads` that one CU holds at once (times the CUs: the limit of a plan with teams)
inline hipError_t tcp_max_blocks(int runs, int threads, int *per_cu)
{
    return runs == 25 ? tcp_max_blocks_25(threads, per_cu) : tcp_max_blocks_9(threads, per_cu);
}
} // namespace tw
