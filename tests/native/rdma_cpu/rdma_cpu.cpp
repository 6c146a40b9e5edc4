// TEST INFRASTRUCTURE: the RDMA slot kernel's launch replaced by the same steps run on the CPU,
// one world after the other and one entity after the other in each slot, so that the unchanged
// host code (rdma_core.cpp, twins_rdma_cli.cpp) and model (rdma_model.h) run as a CPU program
// (with AddressSanitizer: bounds checks of the kernel code) whose results are compared with the
// reference's RDMA fixtures. Never linked into the product.
#include <cstdlib>
#include "../../../twins_amd/csrc/rdma_model.h"
#include "../../../twins_amd/csrc/rdma_kernel.h"

namespace rd
{
static SwWork g_work;      // a switch's slot workspace (LDS on the GPU)
static SwWorkWide g_wide;  // the workspace of a switch of KS + 1 to KW ports (device memory on the GPU)
// TWINS_RDMA_CPU_SERIAL=1: every switch on the serial path of wide switches (switch_step), so that
// the reference's jobs (switches of at most 32 ports) check that path too
static const bool g_serial = getenv("TWINS_RDMA_CPU_SERIAL") != nullptr;
int launch_capacity(int) { return 64; }
// every world of the queue to its end, one after the other (device memory is host memory here)
int launch_worlds(const LaunchCfg *cfg, int, int, hipStream_t)
{
    for (int qi = 0; qi < cfg->nq; qi++)
    {
        const int qe = cfg->queue[qi], w = qe & ((1 << QSHIFT) - 1);
        const Sim &s = *cfg->sims[qe >> QSHIFT];
        const World &wd = s.wd[w];
        WorldRun &r = s.run[w];
        uint32_t *pend = s.pend + wd.pend_off;
        while (!r.finished)
        {
            unsigned act = 0;
            unsigned long long wake = ~0ull;
            for (int e = 0; e < wd.nh + wd.ns; e++)
            {
                if (e < wd.nh)
                {
                    int64_t hw = INT64_MAX;
                    act |= host_step(s, wd.h0 + e, r.slot, &r.flows_done, pend, wd.pend_n, &hw) ? 1u : 0u;
                    if (hw != INT64_MAX && (unsigned long long)hw < wake)
                        wake = (unsigned long long)hw;
                }
                else
                {
                    const int x = wd.s0 + (e - wd.nh);
                    bool busy = false;
                    if (s.sw_busy[x] || switch_has_input(s, x, r.slot))
                        busy = g_serial              ? switch_step(s, x, r.slot, pend, wd.pend_n)
                               : s.si[x].deg <= KS ? switch_slot(s, x, r.slot, pend, wd.pend_n, g_work, 0, 1)
                               : s.si[x].deg <= KW ? switch_slot(s, x, r.slot, pend, wd.pend_n, g_wide, 0, 1)
                                                   : switch_step(s, x, r.slot, pend, wd.pend_n);
                    s.sw_busy[x] = busy ? 1 : 0;
                    act |= busy ? 1u : 0u;
                }
            }
            slot_end(s, wd, r, pend, act, wake);
        }
    }
    return 0;
}
} // namespace rd
