// The RDMA slot kernel's launch, between the host side (rdma_core.cpp) and the kernel
// (rdma_kernel.hip; a CPU build of the same steps replaces it in the CPU check).
#pragma once
#include <hip/hip_runtime_api.h>
#include "rdma_model.h"

namespace rd
{
// a team's barrier and its world's slot decisions (device memory, zeroed before a launch)
struct TeamState
{
    unsigned arrive, gen;
    int world, err;
    unsigned act;
    int stop;
    unsigned long long wake;
};
// one launch: teams of T workgroups take worlds queue[0 .. nq) in order (qnext: the next index);
// an entry is sim << QSHIFT | world: world `world` of the batch whose Sim is sims[sim] (a launch can
// run the worlds of several separately packed batches: the serving loop's requests)
constexpr int QSHIFT = 20;
constexpr int LAUNCH_WAVES = 4; // waves of a launch's 256-thread workgroups
struct LaunchCfg
{
    int T, nq, hpw; // hpw: hosts of a host item (side by side on a wave's lanes, at most 64)
    const int *queue;
    unsigned *qnext;
    TeamState *teams;
    const Sim *const *sims;
    SwWorkWide *wide; // a device-memory switch workspace per wave of the launch (null: no switch of KS + 1 to KW ports)
};
// the workgroups of 256 threads the GPU holds at once (-1: unknown)
int launch_capacity(int threads);
// one launch of nblocks workgroups (teams of cfg->T); cfg_dev is a device copy; 0 or a hipError_t
int launch_worlds(const LaunchCfg *cfg_dev, int nblocks, int threads, hipStream_t st);
} // namespace rd
