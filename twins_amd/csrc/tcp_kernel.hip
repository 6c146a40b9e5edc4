// The TCP slot kernels for gfx950.
//
// tcp_worlds_kernel: a persistent world-queue kernel. Each workgroup takes the next world from
// the batch's queue and runs its slots one after another until the world finishes or the
// launch's slot budget is spent; the slot barrier is the workgroup barrier, so a world's whole
// state is touched by one CU and no grid-wide synchronisation exists. Within a slot, lanes run
// hosts (one serial event loop per lane) and wavefronts run switches (one lane per port of
// 64/k switches), with no barrier between the two: both read only the two previous slots' link
// messages and write their own links' message for this slot (reference: src/kernel.cuh:5-113).
// The last step of a slot is the ticket (one lane): finish test and the exact idle-slot skip.
#include "tcp_model.h"
#include "tcp_kernel.h"

namespace tw
{

template <int RUNS>
__global__ void __launch_bounds__(1024) tcp_worlds_kernel(const BatchT<RUNS> *__restrict__ Bp, int64_t slot_budget)
{
    typedef Model<RUNS> Md;
    const BatchT<RUNS> &B = *Bp;
    extern __shared__ int lds[]; // per wave: rows[64 * 33], in_win[64 * 3], base[64], cnt[64]
    __shared__ int s_qi;
    __shared__ int64_t s_t;
    __shared__ int s_done;
    __shared__ unsigned s_act;
    __shared__ unsigned long long s_wake;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwaves = blockDim.x >> 6;
    constexpr int LDS_PER_WAVE = 64 * 33 + 64 * 3 + 64 + 64;
    int *wl = lds + wv * LDS_PER_WAVE;
    const int gwave = blockIdx.x * nwaves + wv; // this wave's scratch
    int64_t *scr_ts = B.scr_ts + (size_t)gwave * 64 * 2 * MAX_TX;
    uint32_t *scr_ref = B.scr_ref + (size_t)gwave * 64 * 2 * MAX_TX;
    uint8_t *scr_up = B.scr_up + (size_t)gwave * 64 * 2 * MAX_TX;
    for (;;)
    {
        if (tid == 0)
            s_qi = atomicAdd(B.queue_next, 1);
        __syncthreads();
        const int qi = s_qi;
        if (qi >= B.nworlds)
            return;
        const int w = B.world_order[qi];
        const WorldDesc &W = B.worlds[w];
        WorldRun *R = B.run + w;
        const int k = W.k, spw = 64 / k, ngroups = (W.S + spw - 1) / spw;
        for (int64_t it = 0;; it++)
        {
            if (tid == 0)
            {
                s_t = R->slot;
                s_done = R->finished || it >= slot_budget;
                s_act = 0;
                s_wake = ~0ull;
            }
            __syncthreads();
            if (s_done)
                break;
            const int64_t t = s_t;
            unsigned act = 0;
            unsigned long long wk = ~0ull;
            for (int hl = tid; hl < W.H; hl += blockDim.x)
            {
                int64_t x = INT64_MAX;
                act += Md::host_step(B, W, R, hl, t, &x);
                wk = (unsigned long long)x < wk ? (unsigned long long)x : wk;
            }
            for (int g = wv; g < ngroups; g += nwaves)
            {
                const int sl = lane / k, base = (sl < spw ? sl : 0) * k;
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
                act += Md::switch_step(B, W, R, x, t, L);
            }
            if (act)
                atomicAdd(&s_act, act);
            if (wk != ~0ull)
                atomicMin(&s_wake, wk);
            __syncthreads();
            if (tid == 0)
            {
                // the ticket (src/kernel.cuh:79-112)
                unsigned cur = s_act + R->act_cur, prev = R->act_prev;
                bool fin = atomicAdd(&R->flows_done, 0u) >= (unsigned)W.M || t + 1 >= W.max_slots; // (atomics bypass L1)
                int64_t next = t + 1;
                unsigned long long nw = s_wake;
                if (cur == 0 && prev == 0 && nw != ~0ull && (int64_t)nw > next)
                    next = (int64_t)nw < W.max_slots ? (int64_t)nw : W.max_slots;
                R->act_prev = cur;
                R->act_cur = 0;
                R->slot = next;
                R->slots_run++;
                if (fin)
                    R->finished = 1;
            }
        }
    }
}

template <int RUNS>
static hipError_t launch(const void *batch_dev, int blocks, int threads, int64_t slot_budget, hipStream_t stream)
{
    size_t lds = (size_t)(threads / 64) * (64 * 33 + 64 * 3 + 64 + 64) * sizeof(int);
    hipLaunchKernelGGL(tcp_worlds_kernel<RUNS>, dim3(blocks), dim3(threads), lds, stream,
                       reinterpret_cast<const BatchT<RUNS> *>(batch_dev), slot_budget);
    return hipGetLastError();
}

hipError_t tcp_launch(int runs, const void *batch_dev, int blocks, int threads, int64_t slot_budget, hipStream_t stream)
{
    return runs == 25 ? launch<25>(batch_dev, blocks, threads, slot_budget, stream)
                      : launch<9>(batch_dev, blocks, threads, slot_budget, stream);
}

} // namespace tw
