/* twins.h -- the C ABI of libtwins_core.so, the MI355X (gfx950) packet-simulation core of the
 * TwiNS TCP model (`twins`).
 *
 * The reference has no library interface: its boundary is the `twins` command line, the flow /
 * topology / worlds file formats and the fct.csv output, driven by one `main` (src/main.cu). Each
 * entry point below replaces one stage of that `main`; the `twins` binary built here calls them,
 * and so can any host program (ctypes, cgo, JNI: see INTEGRATION.md).
 *
 * Conventions: every function returns 0 on success and a negative code on error, with a message
 * from tw_last_error() (per thread). Handles own their memory; caller buffers are copied and never
 * retained. A batch owns its device memory and one HIP stream; use one host thread per batch.
 * Times are nanoseconds at the boundary (the model runs in integer picoseconds).
 */
#ifndef TWINS_H
#define TWINS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TW_OK = 0, TW_EINVAL = -1, TW_EIO = -2, TW_ENOMEM = -3, TW_EDEVICE = -4, TW_ESTATE = -5 };
enum { TW_TOPO_FATTREE = 0, TW_TOPO_LEAFSPINE = 1, TW_TOPO_GRAPH = 2 };
enum { TW_CC_NEWRENO = 0, TW_CC_DCTCP = 1, TW_CC_CUBIC = 2 };

typedef struct tw_world tw_world;
typedef struct tw_batch tw_batch;

/* One simulation ("world"): topology, flows and parameters, as host arrays.
 * Replaces the set-up part of the reference's main (src/main.cu:223-1015). */
typedef struct tw_world_desc
{
    int32_t topo;                  /* TW_TOPO_* */
    int32_t k;                     /* FatTree: ports per switch (4..32, a power of two) */
    int32_t leaves, spines, hpl;   /* leaf-spine: leaves, spines, hosts per leaf */
    /* topology file (TW_TOPO_GRAPH): hosts, switches and the links in file order; an endpoint is
     * a host h (>= 0) or switch j (encoded -1 - j); a host is the first endpoint of its one link */
    int32_t n_hosts, n_switches, n_links;
    const int32_t *link_a, *link_b;
    /* link failures (TW_TOPO_GRAPH): at fail_ns[i] the first link between switches fail_a[i] and
     * fail_b[i] goes down (fail_up[i] = 0) or up again (1) */
    int32_t n_fail;
    const int64_t *fail_ns;
    const int32_t *fail_a, *fail_b;
    const uint8_t *fail_up;
    /* flows ("messages") in file order: hosts, bytes, start (ns), connection id (-1: its own
     * connection; equal ids share one connection, in order) */
    int32_t n_flows;
    const int32_t *src, *dst;
    const int64_t *size, *start_ns, *conn;
    /* optional ns-3 ECMP hashes per flow (the reference's --ecmp_from); NULL: computed */
    const uint32_t *hash_data, *hash_ack;
    int32_t cc;                    /* TW_CC_* */
    int32_t ecn_k, queue;          /* RED threshold K and queue limit, packets (0: 30, 200) */
    int32_t fin;                   /* senders close their connections after the last flow */
    int32_t spray;                 /* per-packet ECMP (MRG32k3a), with rng_seed / rng_run */
    uint32_t rng_seed;
    uint64_t rng_run;
    int64_t max_slots;             /* 0: 2,000,000 (2 s simulated) */
    int32_t rtt_ring, ooo_segs;    /* per-connection caps (0: 65536, 512) */
} tw_world_desc;

/* Builds a world: routes per failure epoch, connections, ns-3 ECMP hashes and source ports,
 * the per-host application order. (src/main.cu:327-806, 870-963) */
int tw_world_create(const tw_world_desc *desc, tw_world **out);

/* The same from `twins` command-line arguments (flow and topology files; no --worlds):
 * --k --flows --tcp --ecnK --queue --fin --topo --topofile --fail --leaves --spines --hpl --spray
 * --RngSeed --RngRun --max_slots --ecmp_from --rtt_ring --ooo_segs (src/main.cu:155-222) */
int tw_world_create_args(int argc, const char *const *argv, tw_world **out);

typedef struct tw_world_info
{
    int32_t hosts, switches, ports_per_switch, links, conns, flows;
    uint64_t device_bytes; /* device memory this world needs in a batch */
} tw_world_info;
int tw_world_info_get(const tw_world *w, tw_world_info *info);
/* For a world made with given ECMP hashes (--ecmp_from=F, ns-3's hash_data / hash_ack columns):
 * the connections whose hashes computed here (ns-3's Hash32 over the 5-tuple) equal F's, of all
 * connections -- the reference's --hash_check=1 (src/main.cu:723-768). */
int tw_world_hash_check(const tw_world *w, int32_t *same, int32_t *conns);
/* A topology-file world's routes in the format of the ns-3 scenario's DUMP_ROUTES: per switch and
 * host, "route NODE HOST IF..." (ns-3 node number, the output interfaces of the equal-cost routes
 * before any link failure, numbered from 1) -- the reference's --dump_routes=F (src/main.cu:466-486). */
int tw_world_dump_routes(const tw_world *w, const char *path);
int tw_world_destroy(tw_world *w);

/* An explicit launch plan (the batched sweep's autotuner writes one). Team j has team_blocks[j]
 * workgroups and takes worlds from queue team_queue[j]; world i waits in queue world_queue[i]; a
 * queue's worlds are taken in the order they appear in world_order (a permutation of 0..n-1;
 * NULL: batch order). No counterpart in the reference, whose sweep (--worlds) runs every world in
 * one fixed launch shape (src/main.cu:807-1015). */
typedef struct tw_plan
{
    int32_t nteams, nqueues;
    const int32_t *team_blocks;  /* [nteams] */
    const int32_t *team_queue;   /* [nteams] */
    const int32_t *world_queue;  /* [nworlds] */
    const int32_t *world_order;  /* [nworlds] or NULL */
} tw_plan;

/* Packing of worlds into one device batch and the launch shape (0: automatic). */
typedef struct tw_pack_cfg
{
    int32_t chain_runs;      /* recorded event-order runs: 9 (`twins`) or 25 (`twins128`); 0: 9 */
    int32_t block_threads;   /* threads per workgroup (a multiple of 64, <= 256; 0: 256, one workgroup per CU) */
    int32_t blocks;          /* persistent workgroups (a multiple of team_blocks) */
    int32_t team_blocks;     /* workgroups that run one world together (1: a world per workgroup) */
    int32_t hosts_per_wave;  /* hosts a wavefront runs side by side (1..64) */
    int32_t no_graph;        /* 1: launch directly, not through the captured HIP graph */
    int64_t slot_budget;     /* slots a world may run per kernel launch */
    const tw_plan *plan;     /* teams and queues (NULL: one queue, teams from the fields above) */
} tw_pack_cfg;

/* Workgroups of block_threads (0: 256) that the GPU holds at once: the most a plan with teams of
 * several workgroups may use. */
int tw_launch_capacity(int chain_runs, int block_threads, int32_t *resident_blocks);

/* Packs worlds (copied; the handles stay the caller's) into device memory.
 * (src/main.cu:267-326, 807-1015) */
int tw_batch_create(tw_world *const *worlds, int n, const tw_pack_cfg *cfg, tw_batch **out);

typedef struct tw_run_stats
{
    double wall_s;           /* host wall time of the run (the event loop; set-up excluded) */
    double kernel_s;         /* summed kernel time (HIP events) */
    int64_t launches;
    int64_t slots_max;       /* largest per-world slot count executed */
    uint64_t packet_events;  /* frames started on links, all worlds */
    uint64_t flows_done;
    int32_t worlds_done;
    int32_t graph;           /* 1: the launches went through a captured HIP graph */
} tw_run_stats;

/* Runs every world of the batch to completion (all flows done or max_slots).
 * (src/main.cu:1024-1048, the graph-launch loop) */
int tw_batch_run(tw_batch *b, tw_run_stats *stats);
/* tw_batch_run without allocating anything: no hipMalloc or hipFree, no graph, no allocating
 * library call. The launch's scratch memory (the batch record the kernel reads, the teams' barrier
 * words, the world-queue counters and the switch groups' sort space) comes from a static device
 * array of the library's own, 64 MiB per GPU, held until the process ends; a batch whose launch
 * needs more (above about 33 workgroups of 256 threads: make it with tw_pack_cfg.blocks <= 32) is
 * refused with TW_EINVAL. Results equal tw_batch_run's. Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (from several threads or batches at once) race. No
 * counterpart in the reference; it is for callers that cannot allocate device memory while they
 * run (code under stream capture, tight serving loops, harnesses that forbid allocation). */
int tw_batch_run_noalloc(tw_batch *b, tw_run_stats *stats);

/* Per flow of world `world`, in its file order: when the socket got it (established_ns) and when
 * its last byte was acknowledged (completed_ns); -1 if not reached. (src/main.cu:1143-1153) */
int tw_fct_get(tw_batch *b, int world, int64_t *established_ns, int64_t *completed_ns);

/* The world's counters in the reference's numbering: drops, marks, retx, rto, tx-cap,
 * active/ooo/due overflows, retx by state, max in-flight items, rtt/ring overflows, undecided
 * ties (25), packet-events (26); executed slots in [31]. (src/main.cu:1054-1075) */
int tw_counters_get(tw_batch *b, int world, uint64_t ev[32]);

/* The queue trace (the reference's --qtrace=F --qtrace_from=T0 --qtrace_until=T1; src/main.cu:
 * 169-174, 985-994): every enqueue at a host or switch queue at T0 <= t < T1 (ps), the first
 * max_records of them (0: 20,000,000). Call before tw_batch_run; T1 = 0 records nothing, as in
 * the reference. */
int tw_qtrace_enable(tw_batch *b, int64_t from_ps, int64_t until_ps, uint64_t max_records);
/* Writes the trace after tw_batch_run in the reference's format (src/main.cu:1077-1100): one line
 * "ts_ps node dev src dst kind seq-or-ack qlen" per enqueue (ns-3 node and device numbers, kind
 * D / A / S / SA, qlen the queue-disc length after it), stable-sorted by time, node and device,
 * so that each queue's records keep their order (every world's records, by world at equal keys).
 * *records: the records written. */
int tw_qtrace_write(tw_batch *b, const char *path, uint64_t *records);

/* One world's run: slots executed, the device time its team spent on it (seconds of the GPU's
 * constant clock, summed over launches), the team's workgroups, whether it finished, and its
 * packet-events. No counterpart in the reference (the autotuner's measurements). */
typedef struct tw_world_stats
{
    int64_t slots;
    double device_s;
    int32_t team_blocks, finished;
    uint64_t packet_events;
} tw_world_stats;
int tw_world_stats_get(tw_batch *b, int world, tw_world_stats *st);

int tw_batch_destroy(tw_batch *b);

/* The serving loop: one persistent kernel serves requests (one world each) as they arrive.
 * A request goes to the kernel's queue for its size (teams of 1, mid_size or large_size workgroups). It is
 * admitted in arrival order within that queue once its device footprint fits the server's arena
 * and the queue holds fewer than 4 resident requests per team; a request waiting for either
 * does not hold up the other queues. Admitted, it is packed into the arena and stays resident
 * until its result is taken; requests run side by side, and each one's results are those of the
 * same world run alone (tw_batch_run). No counterpart in the reference, which runs one
 * simulation per process (src/main.cu). */
typedef struct tw_server tw_server;
typedef struct tw_server_cfg
{
    int32_t chain_runs;       /* 9 or 25 (0: 9) */
    int32_t block_threads;    /* 0: 256 */
    int32_t small_teams;      /* teams of one workgroup (0: what the GPU holds after the others and a free share) */
    int32_t mid_teams;        /* mid teams (mid_size workgroups), for worlds of 513-4096 hosts and switches (0: 8) */
    int32_t large_teams;      /* large teams (large_size workgroups), for larger worlds (0: 4) */
    int32_t max_requests;     /* requests over the server's life (0: 65536) */
    uint64_t arena_bytes;     /* device memory for resident requests (0: 60% of the free memory) */
    double idle_s;            /* a kernel team idle this long leaves; the kernel is relaunched on demand (0: 2 s) */
    int32_t mid_size;         /* workgroups of a mid team (0: 4) */
    int32_t large_size;       /* workgroups of a large team (0: 16) */
} tw_server_cfg;
enum { TW_REQ_WAITING = 0, TW_REQ_RESIDENT = 1, TW_REQ_DONE = 2, TW_REQ_FAILED = 3, TW_REQ_TAKEN = 4 };
typedef struct tw_server_stats
{
    int32_t waiting, resident, done, failed;
    uint64_t arena_used, arena_bytes, relaunches;
} tw_server_stats;
int tw_server_create(const tw_server_cfg *cfg, tw_server **out);
/* Queues a request (the world is copied); admits what fits. */
int tw_server_submit(tw_server *s, const tw_world *w, int32_t *id);
/* Collects completions, admits waiting requests, relaunches the kernel if it left. */
int tw_server_poll(tw_server *s, tw_server_stats *st);
int tw_server_state(tw_server *s, int32_t id, int32_t *state);
/* A finished request's flow times (as tw_fct_get) and counters (as tw_counters_get); frees its
 * device memory. */
int tw_server_result(tw_server *s, int32_t id, int64_t *established_ns, int64_t *completed_ns, uint64_t ev[32]);
int tw_server_destroy(tw_server *s);
const char *tw_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
