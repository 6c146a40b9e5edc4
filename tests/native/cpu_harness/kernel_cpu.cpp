// TEST INFRASTRUCTURE: runs the device functions of twins_amd/csrc/tcp_model.h on the CPU, in the
// order the gfx950 kernel does (tcp_kernel.hip: per slot every host, every switch's lanes phase
// by phase, then the ticket), over the same batch layout, so that AddressSanitizer can check
// every memory access of the model code and its results can be compared with the oracle's.
// The phases a wavefront separates with wave_sync() run lane after lane here.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

template <typename T, typename U>
static inline T atomicAdd(T *p, U v)
{
    T o = *p;
    *p = o + (T)v;
    return o;
}
template <typename T, typename U>
static inline T atomicMax(T *p, U v)
{
    T o = *p;
    if ((T)v > o)
        *p = (T)v;
    return o;
}
template <typename T, typename U>
static inline T atomicMin(T *p, U v)
{
    T o = *p;
    if ((T)v < o)
        *p = (T)v;
    return o;
}
#define __builtin_amdgcn_fence(a, b) ((void)0)
#define __builtin_amdgcn_wave_barrier() ((void)0)

#include "../../../twins_amd/csrc/tcp_kernel.h"
#include "../../../twins_amd/csrc/tcp_model.h"

namespace tw
{
template <int RUNS>
static void run_cpu(BatchT<RUNS> &B, int64_t budget)
{
    typedef Model<RUNS> Md;
    std::vector<int> lds(Md::SW_LDS_TOTAL);
    for (int wq = 0; wq < B.nwq; wq++)
    for (int qi = B.queue_next[wq]; qi < B.wq_len[wq]; qi = ++B.queue_next[wq])
    {
        const int w = B.world_order[B.wq_off[wq] + qi];
        const WorldDesc &W = B.worlds[w];
        WorldRun *R = B.run + w;
        const int k = W.k;
        for (int64_t it = 0; !R->finished && it < budget; it++)
        {
            const int64_t t = R->slot;
            unsigned act = 0;
            unsigned long long wk = ~0ull;
            for (int hl = 0; hl < W.H; hl++)
            {
                int64_t x = INT64_MAX;
                SndState ls;
                RcvState lr;
                ConnStatic lc;
                act += Md::host_step(B, W, R, hl, t, &x, &ls, &lr, &lc);
                wk = std::min(wk, (unsigned long long)x);
            }
            for (int sw = 0; sw < W.S; sw++)
            {
                std::vector<typename Md::SwLane> lanes(k);
                typename Md::SwLds L;
                L.rows = lds.data();
                L.in_win = lds.data() + 64 * 33;
                L.base = lds.data() + 64 * 33 + 64 * 3;
                L.cnt = lds.data() + 64 * 33 + 64 * 3 + 64;
                for (int q = 0; q < k; q++)
                {
                    auto &x = lanes[q];
                    x.swl = sw;
                    x.q = q;
                    x.active = true;
                    x.lid = W.link0 + W.H + sw * k + q;
                    x.scr_ts = B.scr_ts;
                    x.scr_ref = B.scr_ref;
                    x.scr_up = B.scr_up;
                }
                L.wave = lds.data();
                for (auto &x : lanes)
                    Md::sw_window(B, W, x, t, L);
                if (W.rng0 >= 0)
                {
                    // the lockstep merge, phase by phase over the lanes
                    std::vector<typename Md::Spray> sp(k);
                    for (int q = 0; q < k; q++)
                        Md::spray_init(B, W, lanes[q], t, L, sp[q]);
                    for (;;)
                    {
                        for (int q = 0; q < k; q++)
                            Md::spray_put(lanes[q], L, sp[q]);
                        bool done = true;
                        for (int q = 0; q < k; q++)
                            done &= Md::spray_pick(B, W, R, lanes[q], t, L, sp[q]);
                        if (done)
                            break;
                    }
                    for (int q = 0; q < k; q++)
                        Md::spray_end(B, W, lanes[q], L, sp[q]);
                }
                for (auto &x : lanes)
                    Md::sw_count(B, W, R, x, t, L);
                for (auto &x : lanes)
                    Md::sw_prefix(W, x, L);
                for (auto &x : lanes)
                    Md::sw_base(x, L);
                for (auto &x : lanes)
                    Md::sw_scatter(B, W, x, t, L);
                for (auto &x : lanes)
                    act += Md::sw_egress(B, W, R, x, t, L);
            }
            unsigned cur = act + R->act_cur, prev = R->act_prev;
            bool fin = R->flows_done >= (unsigned)W.M || t + 1 >= W.max_slots;
            int64_t next = t + 1;
            if (cur == 0 && prev == 0 && wk != ~0ull && (int64_t)wk > next)
                next = (int64_t)wk < W.max_slots ? (int64_t)wk : W.max_slots;
            R->act_prev = cur;
            R->act_cur = 0;
            R->slot = next;
            R->slots_run++;
            if (fin)
                R->finished = 1;
        }
        R->team = 1;
    }
}

hipError_t tcp_launch_9(const void *batch_dev, const LaunchShape &s, hipStream_t)
{
    run_cpu(*(BatchT<9> *)batch_dev, s.slot_budget);
    return hipSuccess;
}
hipError_t tcp_launch_25(const void *batch_dev, const LaunchShape &s, hipStream_t)
{
    run_cpu(*(BatchT<25> *)batch_dev, s.slot_budget);
    return hipSuccess;
}
hipError_t tcp_reset(uint32_t *queue_next, int nq, void *teams, size_t team_bytes, hipStream_t)
{
    std::fill(queue_next, queue_next + nq, 0u);
    std::fill((char *)teams, (char *)teams + team_bytes, (char)0);
    return hipSuccess;
}
// the static scratch array of tw_batch_run_noalloc: host memory here
hipError_t tcp_static_scratch(void **ptr)
{
    static std::vector<unsigned char> host(TW_SCRATCH_BYTES);
    *ptr = host.data();
    return hipSuccess;
}
hipError_t tcp_server_launch_9(const void *, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t tcp_server_launch_25(const void *, int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t tcp_max_blocks_9(int, int *per_cu)
{
    *per_cu = 1;
    return hipSuccess;
}
hipError_t tcp_max_blocks_25(int, int *per_cu)
{
    *per_cu = 1;
    return hipSuccess;
}
} // namespace tw
