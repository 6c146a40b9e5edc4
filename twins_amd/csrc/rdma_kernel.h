// The RDMA slot kernel's launch, between the host side (rdma_core.cpp) and the kernel
// (rdma_kernel.hip; a CPU build of the same steps replaces it in the CPU check).
#pragma once
#include <hip/hip_runtime_api.h>
#include "rdma_model.h"

namespace rd
{
// one launch: each world's workgroup runs its slots until it finishes or the launch's time is up
// (sim: the device copy of the Sim); 0 or a hipError_t
int launch_worlds(const Sim *sim, int nworlds, int threads, hipStream_t st);
} // namespace rd
