// The RDMA slot kernel for gfx950. A world runs on a team of T workgroups (256 threads, 4 waves
// each) for all of its slots; teams take worlds from a queue, longest first. In a slot the world's
// work items are dealt over the team's waves: a host item is up to hpw hosts side by side (a lane
// runs one host's events, src/rdma/rstep.cuh:1649-2112), a switch item is one switch on one wave
// (its ports in LDS; a safe slot runs each egress port's events on a lane of its own, otherwise lane
// 0 runs them all; rdma_model.h switch_slot). Entities only read what earlier slots filed
// (src/rdma/rdma.h:12-16), so a slot needs one team barrier; the last workgroup to arrive decides the
// next slot (slot_end: finish, idle skip). A launch ends when the queue is empty or its time (about
// a second) is up; the host launches again for the worlds left.
#include <hip/hip_runtime.h>
#include "rdma_model.h"
#include "rdma_kernel.h"

namespace rd
{
constexpr long long LAUNCH_TICKS = 100000000;          // a launch's time: 1 s of the 100 MHz constant clock
constexpr unsigned long long WAIT_TICKS = 2000000000ull; // a barrier wait gives up after 20 s
constexpr int WAVES = LAUNCH_WAVES;                    // waves of a 256-thread workgroup

// (its LDS holds one workgroup per CU: one wave per SIMD, so all registers may be used)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) rdma_team_kernel(const LaunchCfg *__restrict__ lp)
{
    const LaunchCfg L = *lp;
    __shared__ SwWork work[WAVES];
    __shared__ int64_t s_slot;
    __shared__ int s_w, s_stop, s_err;
    __shared__ unsigned s_act;
    __shared__ unsigned long long s_wake;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int T = L.T;
    // a team's workgroups have consecutive ids, so they spread over the 8 XCDs (workgroups reach
    // the XCDs round-robin by id). Teams kept on one XCD each were measured slower: the reference's
    // 80 short RDMA jobs as one batch 13.0 s against 9.7 s, the 36 medium ones 21.9 s against 18.7 s.
    const int team = blockIdx.x / T, rank = blockIdx.x % T;
    TeamState *ts = L.teams + team;
    const long long t_launch = wall_clock64();
    unsigned gen = 0;

    // the team barrier; the last workgroup to arrive runs `last` first. Within a world (fenced
    // false) it has no cache fences: what the workgroups hand each other goes through coherent
    // accesses (rdma_model.h), so each wave only waits for its own stores to complete, and the
    // entities' own records stay in the caches from slot to slot (an agent-scope fence on the
    // multi-XCD GPU writes back and invalidates the L2). Claiming a world and leaving one are fenced.
    auto barrier = [&](auto last, bool fenced) -> bool {
        __builtin_amdgcn_s_waitcnt(0); // (this wave's stores have completed)
        __syncthreads();
        if (tid == 0)
        {
            if (fenced)
                __threadfence();
            if (T == 1)
                last();
            else
            {
                const unsigned g = gen;
                if (atomicAdd(&ts->arrive, 1u) == (unsigned)T - 1)
                {
                    last();
                    atomicExch(&ts->arrive, 0u);
                    __builtin_amdgcn_s_waitcnt(0); // (the ticket's stores before the release of the others)
                    atomicAdd(&ts->gen, 1u);
                }
                else
                {
                    const unsigned long long w0 = wall_clock64();
                    while (coh_load(&ts->gen) == g)
                    {
                        __builtin_amdgcn_s_sleep(1);
                        if (wall_clock64() - w0 > WAIT_TICKS)
                        {
                            atomicExch(&ts->err, 1);
                            break;
                        }
                    }
                }
                gen = g + 1;
            }
            if (fenced)
                __threadfence();
            s_err = T > 1 && __hip_atomic_load(&ts->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        return s_err;
    };

    for (;;)
    {
        // the team's next world (none once the launch's time is up: the host launches again)
        if (barrier([&] {
                ts->world = wall_clock64() - t_launch > LAUNCH_TICKS ? L.nq : (int)atomicAdd(L.qnext, 1u);
                ts->act = 0;
                ts->wake = ~0ull;
                ts->stop = 0;
            }, true))
            return;
        if (tid == 0)
            s_w = __hip_atomic_load(&ts->world, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (s_w >= L.nq)
            return;
        const int qe = L.queue[s_w], w = qe & ((1 << QSHIFT) - 1);
        const Sim &s = *L.sims[qe >> QSHIFT];
        const World wd = s.wd[w];
        WorldRun &r = s.run[w];
        uint32_t *pend = s.pend + wd.pend_off;
        const int hpw = L.hpw, nhw = (wd.nh + hpw - 1) / hpw, items = nhw + wd.ns, TW = T * WAVES, gw = rank * WAVES + wv;
        const long long t_world = wall_clock64();
        for (;;)
        {
            if (tid == 0)
            {
                s_slot = __hip_atomic_load(&r.slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_stop = __hip_atomic_load(&r.finished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                         __hip_atomic_load(&ts->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_act = 0;
                s_wake = ~0ull;
            }
            __syncthreads();
            if (s_stop)
                break;
            const int64_t sl = s_slot;
            unsigned a = 0;
            unsigned long long wk = ~0ull;
#ifdef RD_PROF
            long long p0 = wall_clock64(), ph = 0, psw = 0;
#endif
            for (int it = gw; it < items; it += TW)
            {
#ifdef RD_PROF
                const long long q0 = wall_clock64();
#endif
                if (it < nhw)
                {
                    const int e = it * hpw + lane;
                    if (lane < hpw && e < wd.nh)
                    {
                        int64_t hw = INT64_MAX;
                        a |= host_step(s, wd.h0 + e, sl, &r.flows_done, pend, wd.pend_n, &hw) ? 1u : 0u;
                        if (hw != INT64_MAX && (unsigned long long)hw < wk)
                            wk = (unsigned long long)hw;
                    }
                }
                else
                {
                    const int x = wd.s0 + (it - nhw);
                    const int deg = s.si[x].deg;
                    bool busy = false;
                    if (s.sw_busy[x] || switch_has_input(s, x, sl))
                    {
                        if (deg <= KS)
                            busy = switch_slot(s, x, sl, pend, wd.pend_n, work[wv], lane, 64);
                        else if (deg <= KW && L.wide)
                            busy = switch_slot(s, x, sl, pend, wd.pend_n, L.wide[(size_t)blockIdx.x * WAVES + wv], lane, 64);
                        else
                        {
                            int b = 0;
                            if (lane == 0)
                                b = switch_step(s, x, sl, pend, wd.pend_n) ? 1 : 0;
                            busy = __shfl(b, 0) != 0;
                        }
                    }
                    if (lane == 0)
                    {
                        s.sw_busy[x] = busy ? 1 : 0;
                        a |= busy ? 1u : 0u;
                    }
                }
#ifdef RD_PROF
                (it < nhw ? ph : psw) += wall_clock64() - q0;
#endif
            }
#ifdef RD_PROF
            // (profiling builds: the waves' host and switch item time, and the time to the slot's end)
            if (lane == 0)
            {
                if (ph)
                    atomicAdd(&s.ev[33], (unsigned long long)ph);
                if (psw)
                    atomicAdd(&s.ev[34], (unsigned long long)psw);
            }
            const long long pb = wall_clock64();
            (void)p0;
#endif
            if (a)
                atomicOr(&s_act, a);
            if (wk != ~0ull)
                atomicMin(&s_wake, wk);
            __syncthreads();
            // this workgroup's share joins the team's before it arrives at the barrier
            if (tid == 0)
            {
                if (s_act)
                    atomicOr(&ts->act, s_act);
                if (s_wake != ~0ull)
                    atomicMin(&ts->wake, s_wake);
            }
            if (barrier([&] {
                    slot_end(s, wd, r, pend, atomicOr(&ts->act, 0u), atomicMin(&ts->wake, ~0ull));
                    coh_store(&ts->act, 0u);
                    coh_store(&ts->wake, ~0ull);
                    if (wall_clock64() - t_launch > LAUNCH_TICKS)
                        coh_store(&ts->stop, 1);
                }, false))
                return;
#ifdef RD_PROF
            if (tid == 0 && rank == 0)
                atomicAdd(&s.ev[35], (unsigned long long)(wall_clock64() - pb));
#endif
        }
        if (tid == 0 && rank == 0)
            r.ticks += wall_clock64() - t_world;
    }
}

int launch_capacity(int threads)
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rdma_team_kernel, threads, 0) != hipSuccess)
        return -1;
    return per_cu * cus;
}

int launch_worlds(const LaunchCfg *cfg_dev, int nblocks, int threads, hipStream_t st)
{
    hipLaunchKernelGGL(rdma_team_kernel, dim3(nblocks), dim3(threads), 0, st, cfg_dev);
    return (int)hipGetLastError();
}
} // namespace rd
