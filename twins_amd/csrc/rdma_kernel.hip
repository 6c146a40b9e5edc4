// The RDMA slot kernel for gfx950: one workgroup per world, persistent over the world's slots.
// In a slot the workgroup's lanes take the world's hosts and switches in turn (a lane runs one
// entity's events; entities only read what earlier slots filed, src/rdma/rdma.h:12-16), the
// workgroup's barrier closes the slot, and lane 0 decides the next one (slot_end: finish, idle skip).
// A launch ends when every world has finished or its time (about a second) is up; the host then
// launches again.
#include <hip/hip_runtime.h>
#include "rdma_model.h"
#include "rdma_kernel.h"

namespace rd
{
constexpr long long LAUNCH_TICKS = 100000000; // 1 s of the 100 MHz constant clock

__global__ void __launch_bounds__(256) rdma_worlds_kernel(const Sim *sp)
{
    const Sim &s = *sp;
    const int w = blockIdx.x;
    const World wd = s.wd[w];
    WorldRun &r = s.run[w];
    uint32_t *pend = s.pend + wd.pend_off;
    __shared__ int64_t slot;
    __shared__ int stop;
    __shared__ unsigned act;
    __shared__ unsigned long long wake;
    const long long t0 = wall_clock64();
    const int ne = wd.nh + wd.ns;
    for (;;)
    {
        if (threadIdx.x == 0)
        {
            slot = r.slot;
            stop = r.finished || wall_clock64() - t0 > LAUNCH_TICKS;
            act = 0;
            wake = ~0ull;
        }
        __syncthreads();
        if (stop)
            break;
        const int64_t sl = slot;
        unsigned a = 0;
        unsigned long long wk = ~0ull;
        for (int e = threadIdx.x; e < ne; e += blockDim.x)
        {
            if (e < wd.nh)
            {
                int64_t hw = INT64_MAX;
                a |= host_step(s, wd.h0 + e, sl, &r.flows_done, pend, wd.pend_n, &hw) ? 1u : 0u;
                if (hw != INT64_MAX && (unsigned long long)hw < wk)
                    wk = (unsigned long long)hw;
            }
            else
            {
                const int x = wd.s0 + (e - wd.nh);
                bool busy = false;
                if (s.sw_busy[x] || switch_has_input(s, x, sl))
                    busy = switch_step(s, x, sl, pend, wd.pend_n);
                s.sw_busy[x] = busy ? 1 : 0;
                a |= busy ? 1u : 0u;
            }
        }
        if (a)
            atomicOr(&act, a);
        if (wk != ~0ull)
            atomicMin(&wake, wk);
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0)
            slot_end(s, wd, r, pend, act, wake);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        r.ticks += wall_clock64() - t0;
}

int launch_worlds(const Sim *sim, int nworlds, int threads, hipStream_t st)
{
    hipLaunchKernelGGL(rdma_worlds_kernel, dim3(nworlds), dim3(threads), 0, st, sim);
    return (int)hipGetLastError();
}
} // namespace rd
