// The TCP slot kernels for gfx950.
//
// tcp_worlds_kernel: a persistent world-queue kernel run by teams of T workgroups. A team takes
// the next world from the batch's queue and runs its slots one after another until the world
// finishes or the launch's slot budget is spent. Within a slot the team's wavefronts share the
// world's entities: "work items" are groups of `hpw` hosts (one serial event loop per lane) and
// groups of 64/k switches (one lane per port), dealt round-robin over the team's waves, with no
// barrier between hosts and switches: both read only the two previous slots' link messages and
// write their own links' message for this slot (reference: src/kernel.cuh:5-113). The slot ends
// with the team barrier, whose last arrival is the ticket (finish test, exact idle-slot skip).
//
// The launch plan gives each team its size T, its workgroups and the queue of worlds it takes
// from, so one launch can run small worlds a workgroup each and large ones over many CUs.
// T = 1 needs only the workgroup barrier, so a world's state stays within one CU. T > 1
// synchronises through an agent-scope arrive counter and generation word per team; the grid is
// at most what the GPU holds at once, so all of a team's workgroups are resident, and every wait
// is bounded (a timeout ends the world with an error flag instead of hanging the queue). Each world's team
// accounts the device clock time it spent on it (WorldRun::ticks), the input of the autotuner.
#include <mutex>

#include "tcp_model.h"
#include "tcp_kernel.h"

namespace tw
{

// per wave, for the item it runs: a switch item's rows, in_win, base, cnt, sort and route slots;
// a host item's slots for the connection records of an event, one SndState, one RcvState and one
// ConnStatic per lane. A wave runs its items one after another, so the two share one region.
constexpr int LDS_SWITCH = Model<9>::SW_LDS_TOTAL;
constexpr int LDS_SND = 0, LDS_RCV = LDS_SND + 64 * (int)(sizeof(SndState) / 4);
constexpr int LDS_CONN = LDS_RCV + 64 * (int)(sizeof(RcvState) / 4);
constexpr int LDS_RECORDS = LDS_CONN + 64 * (int)(sizeof(ConnStatic) / 4);
constexpr int LDS_PER_WAVE = LDS_RECORDS > LDS_SWITCH ? LDS_RECORDS : LDS_SWITCH;
static_assert(sizeof(SndState) % 128 == 0 && sizeof(RcvState) % 128 == 0 && sizeof(ConnStatic) % 16 == 0, "LDS record alignment");
static_assert(4 * (size_t)LDS_PER_WAVE * 4 + 1024 <= 160 * 1024, "a 256-thread workgroup's LDS");
constexpr unsigned long long WAIT_TICKS = 2000000000ull;  // a team-barrier wait gives up after 20 s (100 MHz clock)
constexpr unsigned long long SERVER_MAX_WAIT = (unsigned long long)(SERVER_WAIT_S * 100e6); // (100 MHz clock)

__device__ __forceinline__ unsigned ld_acquire(const unsigned *p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT); }

template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)((uint64_t)hi << 32 | lo);
}

// The world loop of both persistent kernels: SRV = false takes worlds from a batch's queues
// (tcp_worlds_kernel), SRV = true takes requests from the serving loop's queues
// (tcp_server_kernel). One body, specialised at compile time, so the batch kernel carries nothing
// of the server's.
template <int RUNS, bool SRV>
__device__ __forceinline__ void world_loop(const BatchT<RUNS> *__restrict__ Bp, const ServerCtl *__restrict__ Sp, int64_t slot_budget)
{
    typedef Model<RUNS> Md;
    extern __shared__ __attribute__((aligned(128))) int lds[];
    __shared__ int s_w;
    __shared__ int64_t s_t;
    __shared__ int s_done, s_err;
    __shared__ unsigned s_act;
    __shared__ unsigned long long s_wake;
    __shared__ const BatchT<RUNS> *s_b;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwaves = blockDim.x >> 6;
    const int32_t *team_of = SRV ? Sp->team_of : Bp->team_of;
    const TeamPlan *plans = SRV ? Sp->plan : Bp->plan;
    const int team = team_of[blockIdx.x] >> 8;
    const TeamPlan P = plans[team];
    const int T = P.T, rank = team_of[blockIdx.x] & 255;
    const int TW = T * nwaves, gw = rank * nwaves + wv; // the team's waves; this wave's place
    TeamState *ts = (SRV ? Sp->teams : Bp->teams) + team;
    int *wl = lds + wv * LDS_PER_WAVE;
    const int gwave = blockIdx.x * nwaves + wv; // this wave's global scratch
    int64_t *scr_ts = (SRV ? Sp->scr_ts : Bp->scr_ts) + (size_t)gwave * 64 * 2 * MAX_TX;
    uint32_t *scr_ref = (SRV ? Sp->scr_ref : Bp->scr_ref) + (size_t)gwave * 64 * 2 * MAX_TX;
    uint8_t *scr_up = (SRV ? Sp->scr_up : Bp->scr_up) + (size_t)gwave * 64 * 2 * MAX_TX;
    unsigned gen = 0; // the team barrier's generation this block has passed
    // a barrier wait's limit; the server's claim waits as long as a request may take to come
    unsigned long long wait_limit = WAIT_TICKS;

    // The team barrier; the last workgroup to arrive runs `last` first. Fenced (a world's start
    // and end): agent-scope release and acquire, so that the world's records, written with plain
    // stores, are in memory when it ends and no stale copy is read when it starts. Unfenced (the
    // barrier between slots): what workgroups hand each other from slot to slot goes through
    // device-coherent accesses only (coh_load / coh_store, tcp_state.h), so it is enough that
    // every wave's stores have completed (s_waitcnt) before its workgroup arrives; the caches,
    // and each entity's records in them, stay warm from slot to slot.
    auto team_barrier = [&](auto last, bool fenced) -> bool {
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
                unsigned g = gen;
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
                        if (wall_clock64() - w0 > (SRV ? wait_limit : WAIT_TICKS))
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
        int w = 0;
        if constexpr (!SRV)
        {
            // the team's next world, from its queue
            if (team_barrier([&] { ts->world = atomicAdd(Bp->queue_next + P.queue, 1); }, true))
                return;
            if (tid == 0)
                s_w = __hip_atomic_load(&ts->world, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            const int qi = s_w;
            if (qi >= Bp->wq_len[P.queue])
                return;
            w = Bp->world_order[Bp->wq_off[P.queue] + qi];
        }
        else
        {
            // claim the queue's next index, then wait until the host publishes its request: -1
            // when the server shuts down, or when idle and the claim could be given back
            const ServerCtl &S = *Sp;
            const int q = P.queue;
            wait_limit = SERVER_MAX_WAIT + WAIT_TICKS;
            const bool err = team_barrier([&] {
                unsigned i = atomicAdd(S.next + q, 1u);
                int got = -1;
                const unsigned long long w0 = wall_clock64();
                for (;;)
                {
                    if (__hip_atomic_load(S.shutdown, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
                        break;
                    if (i < __hip_atomic_load(S.tail + q, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM))
                    {
                        got = (int)i;
                        break;
                    }
                    const unsigned long long waited = wall_clock64() - w0;
                    if ((waited > S.idle_ticks && atomicCAS(S.next + q, i + 1, i) == i + 1) || waited > SERVER_MAX_WAIT)
                        break;
                    __builtin_amdgcn_s_sleep(8);
                }
                ts->world = got;
            }, true);
            wait_limit = WAIT_TICKS;
            if (err)
                return;
            if (tid == 0)
            {
                s_w = __hip_atomic_load(&ts->world, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (s_w >= 0)
                    s_b = reinterpret_cast<const BatchT<RUNS> *>(S.req[(size_t)q * S.cap + s_w].batch);
            }
            __syncthreads();
            if (s_w < 0)
                return;
        }
        // (the request's batch pointer made wave-uniform, so its fields load into scalar registers
        // as a kernel argument's do)
        const BatchT<RUNS> &B = SRV ? *uniform_ptr(s_b) : *Bp;
        const WorldDesc &W = B.worlds[w];
        int hpw = P.hpw, spw = 64 / W.k;
        if (hpw <= 0)
            slot_items(W.H, W.S, W.k, T * nwaves, &hpw, &spw);
        WorldRun *R = B.run + w;
        const int k = W.k;
        const int host_items = (W.H + hpw - 1) / hpw, items = host_items + (W.S + spw - 1) / spw;
        const unsigned long long c_begin = wall_clock64();
        for (int64_t it = 0;; it++)
        {
            if (tid == 0)
            {
                s_t = __hip_atomic_load(&R->slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_done = __hip_atomic_load(&R->finished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || it >= slot_budget;
                s_act = 0;
                s_wake = ~0ull;
            }
            __syncthreads();
            if (s_done)
            {
                // every workgroup of the team leaves at the same slot, after a fenced barrier (the
                // world's records in memory before its results are read); the first accounts the
                // time (and, serving, tells the host)
                (void)team_barrier([] {}, true);
                if (rank == 0 && tid == 0)
                {
                    R->ticks += wall_clock64() - c_begin;
                    R->team = T;
                    if constexpr (SRV)
                    {
                        __threadfence_system();
                        __hip_atomic_store(Sp->done + Sp->req[(size_t)P.queue * Sp->cap + s_w].id, 1, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                    }
                }
                break;
            }
            const int64_t t = s_t;
            unsigned act = 0;
            unsigned long long wk = ~0ull;
#ifdef TW_PROF
            __shared__ unsigned long long p_host, p_sw;
            const unsigned long long p0 = wall_clock64();
            unsigned long long p_h = 0, p_s = 0;
            if (tid == 0)
                p_host = p_sw = 0;
            __syncthreads();
#endif
            for (int item = gw; item < items; item += TW)
            {
                if (item < host_items)
                {
                    const int hl = item * hpw + lane;
#ifdef TW_PROF
                    const unsigned long long q0 = wall_clock64();
#endif
                    if (lane < hpw && hl < W.H)
                    {
                        int64_t x = INT64_MAX;
                        act += Md::host_step(B, W, R, hl, t, &x, (SndState *)(wl + LDS_SND) + lane, (RcvState *)(wl + LDS_RCV) + lane,
                                             (ConnStatic *)(wl + LDS_CONN) + lane);
                        wk = (unsigned long long)x < wk ? (unsigned long long)x : wk;
                    }
#ifdef TW_PROF
                    p_h += wall_clock64() - q0;
#endif
                    continue;
                }
                const int g = item - host_items, sl = lane / k, base = (sl < spw ? sl : 0) * k;
                typename Md::SwLane x;
                x.swl = g * spw + sl;
                x.q = lane % k;
                x.active = sl < spw && x.swl < W.S;
                x.lid = W.link0 + W.H + x.swl * k + x.q;
                const size_t so = (size_t)base * 2 * MAX_TX;
                x.scr_ts = scr_ts + so;
                x.scr_ref = scr_ref + so;
                x.scr_up = scr_up + so;
                typename Md::SwLds L;
                L.rows = wl + base * (k + 1);
                L.in_win = wl + 64 * 33 + base * 3;
                L.base = wl + 64 * 33 + 64 * 3 + base;
                L.cnt = wl + 64 * 33 + 64 * 3 + 64 + base;
                L.wave = wl;
#ifdef TW_PROF
                const unsigned long long q0 = wall_clock64();
#endif
                act += Md::switch_step(B, W, R, x, t, L);
#ifdef TW_PROF
                p_s += wall_clock64() - q0;
#endif
            }
#ifdef TW_PROF
            atomicMax(&p_host, p_h);
            atomicMax(&p_sw, p_s);
#endif
            if (act)
                atomicAdd(&s_act, act);
            if (wk != ~0ull)
                atomicMin(&s_wake, wk);
            __syncthreads();
            if (tid == 0 && T > 1)
            {
                if (s_act)
                    atomicAdd(&ts->act, s_act);
                if (s_wake != ~0ull)
                    atomicMin(&ts->wake, s_wake);
            }
#ifdef TW_PROF
            if (tid == 0)
            {
                atomicMax(&ts->prof[0], p_host);
                atomicMax(&ts->prof[1], p_sw);
                atomicMax(&ts->prof[2], wall_clock64() - p0);
            }
#endif
            const bool err = team_barrier([&] {
                // the ticket (src/kernel.cuh:79-112): finish, or the next slot with the exact idle skip
                unsigned a = T > 1 ? atomicExch(&ts->act, 0u) : s_act;
                unsigned long long nw = T > 1 ? atomicExch(&ts->wake, ~0ull) : s_wake;
                unsigned cur = a + coh_load(&R->act_cur), prev = coh_load(&R->act_prev);
                bool fin = atomicAdd(&R->flows_done, 0u) >= (unsigned)W.M || t + 1 >= W.max_slots;
                int64_t next = t + 1;
                if (cur == 0 && prev == 0 && nw != ~0ull && (int64_t)nw > next)
                    next = (int64_t)nw < W.max_slots ? (int64_t)nw : W.max_slots;
#ifdef TW_PROF
                {
                    const unsigned long long now = wall_clock64();
                    const unsigned long long before = atomicExch(&ts->prof[3], now);
                    if (before)
                        atomicAdd(&R->ev[EV_PROF_PERIOD], now - before);
                    atomicAdd(&R->ev[EV_PROF_HOSTS], atomicExch(&ts->prof[0], 0ull));
                    atomicAdd(&R->ev[EV_PROF_SWITCHES], atomicExch(&ts->prof[1], 0ull));
                    atomicAdd(&R->ev[EV_PROF_WORK], atomicExch(&ts->prof[2], 0ull));
                }
#endif
                coh_store(&R->act_prev, cur);
                coh_store(&R->act_cur, 0u);
                coh_store(&R->slot, next);
                coh_store(&R->slots_run, coh_load(&R->slots_run) + 1);
                if (fin)
                    coh_store(&R->finished, 1);
            }, false);
            if (err)
            {
                if (tid == 0)
                {
                    coh_store(&R->error, 1);
                    coh_store(&R->finished, 1);
                    if constexpr (SRV)
                        if (rank == 0)
                        {
                            __threadfence_system();
                            __hip_atomic_store(Sp->done + Sp->req[(size_t)P.queue * Sp->cap + s_w].id, 2, __ATOMIC_RELEASE,
                                               __HIP_MEMORY_SCOPE_SYSTEM);
                        }
                }
                return;
            }
        }
    }
}

template <int RUNS>
__global__ void __launch_bounds__(256) tcp_worlds_kernel(const BatchT<RUNS> *__restrict__ Bp, int64_t slot_budget)
{
    world_loop<RUNS, false>(Bp, nullptr, slot_budget);
}

// The serving loop's persistent kernel. Requests are single-world batches already in device
// memory; the host appends them to per-queue rings in pinned host memory and publishes each
// queue's tail there. A team claims the next index of its queue, waits until the host has
// published it (or the server shuts down), runs the world to its end and flags it done in
// pinned memory. Every wait is bounded: an idle team gives its claim back and leaves after
// ServerCtl::idle_ticks, the host's shutdown flag ends every wait, and the host relaunches the
// kernel when it finds it gone.
template <int RUNS>
__global__ void __launch_bounds__(256) tcp_server_kernel(const ServerCtl *__restrict__ Sp)
{
    world_loop<RUNS, true>(nullptr, Sp, INT64_MAX);
}

// This file is compiled once per chain length (TW_RUNS = 9 or 25, two objects built in parallel:
// fully inlined, each takes minutes); tcp_kernel.h dispatches on the run count.
#ifndef TW_RUNS
#define TW_RUNS 9
#endif
#define TW_CAT(a, b) a##b
#define TW_NAME(a, b) TW_CAT(a, b)

hipError_t TW_NAME(tcp_launch_, TW_RUNS)(const void *batch_dev, const LaunchShape &s, hipStream_t stream)
{
    // an ordinary launch: the grid never exceeds the workgroups the GPU holds at once (the plan
    // checks it), so a team's workgroups are resident together, and a team barrier that still
    // waited too long ends the world with an error. No cooperative launch: it takes a device-wide
    // resource that one process can hold while another waits for it.
    const size_t lds = (size_t)(s.threads / 64) * LDS_PER_WAVE * sizeof(int);
    const BatchT<TW_RUNS> *b = reinterpret_cast<const BatchT<TW_RUNS> *>(batch_dev);
    hipLaunchKernelGGL(tcp_worlds_kernel<TW_RUNS>, dim3(s.blocks), dim3(s.threads), lds, stream, b, s.slot_budget);
    return hipGetLastError();
}

hipError_t TW_NAME(tcp_server_launch_, TW_RUNS)(const void *ctl_dev, int blocks, int threads, hipStream_t stream)
{
    const size_t lds = (size_t)(threads / 64) * LDS_PER_WAVE * sizeof(int);
    const ServerCtl *c = reinterpret_cast<const ServerCtl *>(ctl_dev);
    hipLaunchKernelGGL(tcp_server_kernel<TW_RUNS>, dim3(blocks), dim3(threads), lds, stream, c);
    return hipGetLastError();
}

hipError_t TW_NAME(tcp_max_blocks_, TW_RUNS)(int threads, int *per_cu)
{
    const size_t lds = (size_t)(threads / 64) * LDS_PER_WAVE * sizeof(int);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, tcp_worlds_kernel<TW_RUNS>, threads, lds);
}

#if TW_RUNS == 9
// clears the world-queue counters and the teams' barrier words before a launch (a kernel, not a
// memset, so that the pair can be captured in a graph)
__global__ void tcp_reset_kernel(uint32_t *__restrict__ a, int na, uint32_t *__restrict__ b, int nb)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += gridDim.x * blockDim.x)
    {
        if (i < na)
            a[i] = 0;
        else
            b[i - na] = 0;
    }
}

hipError_t tcp_reset(uint32_t *queue_next, int nq, void *teams, size_t team_bytes, hipStream_t stream)
{
    const int nb = (int)(team_bytes / 4);
    int blocks = (nq + nb + 255) / 256;
    blocks = blocks < 1 ? 1 : blocks > 1024 ? 1024 : blocks;
    hipLaunchKernelGGL(tcp_reset_kernel, dim3(blocks), dim3(256), 0, stream, queue_next, nq, (uint32_t *)teams, nb);
    return hipGetLastError();
}

// The scratch memory of tw_batch_run_noalloc: a launch's batch record, team barrier words, queue
// counters and switch sort space, in a device array of the library's own (every GPU the library
// runs on holds one copy until the process ends), so that the entry point allocates nothing. Host
// code never passes the array's name to a kernel: the address to use is the one
// hipGetSymbolAddress gives, looked up once per device with that device current.
static __device__ __align__(256) unsigned char scratch[TW_SCRATCH_BYTES];

hipError_t tcp_static_scratch(void **ptr)
{
    static std::mutex mu;
    static void *per_dev[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= 64)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!per_dev[dev])
    {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        per_dev[dev] = p;
    }
    *ptr = per_dev[dev];
    return hipSuccess;
}
#endif

} // namespace tw
