// TEST INFRASTRUCTURE, built only into oracle/_ref (never into the product): the driver with which
// the reference's own RDMA model (src/rdma/rmain.cu, rstep.cuh, rdma.h, compiled unchanged for the
// CPU with TWINS_CPU) runs plain flow files. The reference's CPU build replaces its GPU launch loop
// by island_run() from src/rdma/island.h, whose driver steps NVLink DAG islands only
// (src/rdma/island.h:160-166, 225-236; island_setup runs for DAG worlds only, src/rdma/rmain.cu:1686).
// This file takes that header's place on the include path and steps every entity of the world
// slot by slot exactly as Island::step does (src/rdma/island.h:156-166): one lane per block,
// blocks in order, the last block's slot_end advancing the slot, until the finish flag.
#pragma once

static string g_isl_in, g_isl_map, g_isl_out, g_isl_shm, g_isl_rings;
static int g_isl_side = 0, g_isl_nrings = 0, g_isl_ringlog = 22;

namespace rdma
{
uint64_t g_tail_side = 0;
// one process, no islands: every tail node is local
TailNode island_remote(uint64_t) { return TailNode{}; }
} // namespace rdma

static void island_world(World &)
{
    fprintf(stderr, "oracle/_ref driver: DAG flow files are not supported\n");
    exit(1);
}
static void island_setup(Sim &) {}

static void island_run(Sim &s, int64_t)
{
    if (s.ns_b || s.nm)
    {
        fprintf(stderr, "oracle/_ref driver: switches wider than %d ports and multi-NIC hosts are not supported\n", s.ks);
        exit(1);
    }
    blockDim.x = 1;
    threadIdx.x = 0;
    std::vector<char> smem(1 << 20);
    const int nb = s.n1 + s.ns_s;
    while (!*s.finished)
    {
        const int64_t slot = *s.slot;
        for (int b = 0; b < nb; b++)
        {
            blockIdx.x = b;
            slot_entity<false>(s, b, s.bi_s[b], slot, smem.data(), 1);
        }
    }
}
