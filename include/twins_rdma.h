/* twins_rdma.h -- the C ABI of the RDMA model (`twins-rdma`) in libtwins_core.so: hosts with RDMA
 * NICs (queue pairs paced by rate and window, go-back-N with NACKs), switches with a shared-buffer
 * MMU (PFC with dynamic thresholds and headroom, ECN marking at dequeue), congestion control DCQCN,
 * HPCC or DCTCP over RDMA; flow completion times equal to the HPCC simulator's.
 *
 * The reference has no library interface: `twins-rdma` (src/rdma/rmain.cu) reads an HPCC config
 * file, a topology file and a flow file (or a job file of such worlds), runs them and writes
 * "id start_ns fct_ns" per finished flow. Each entry point below replaces one stage of that main;
 * the `twins-rdma` binary built here calls them.
 *
 * Conventions as twins.h: 0 on success, a negative TW_E* code on error with a message from
 * trd_last_error() (per thread); handles own their memory; caller buffers are copied. Times are ns.
 */
#ifndef TWINS_RDMA_H
#define TWINS_RDMA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct trd_world trd_world;
typedef struct trd_batch trd_batch;

/* One world: an HPCC config file (`KEY value ...` lines, rdma-ref's), overrides `KEY=VALUE` (or
 * KEY=V1,V2,... for a key with several values), a TwiNS topology file (`hosts H`, `switches S`,
 * `link A B` with A, B as hN or sN) and a flow file (a header line, then `src,dst,bytes,start_ns`).
 * Every link has rate_bps and delay_ns (0: 100 Gb/s, 1000 ns).
 * Replaces build_world (src/rdma/rmain.cu:364-1021) with read_config / set_override (111-143). */
int trd_world_create(const char *config_path, const char *const *overrides, int32_t n_overrides, const char *topo_path,
                     const char *flow_path, uint64_t rate_bps, int64_t delay_ns, trd_world **out);
/* its size: hosts, switches, flows; and maxRtt / maxBdp (the reference prints them) */
int trd_world_info(const trd_world *w, int32_t *hosts, int32_t *switches, int32_t *flows, uint64_t *max_rtt_ns, uint32_t *max_bdp);
void trd_world_destroy(trd_world *w);

typedef struct trd_pack_cfg
{
    int32_t q3cap;       /* frames a switch port's data queue holds (0: 4096; src/rdma/rmain.cu:1037) */
    int32_t tail_log2;   /* tail_exact 0: a world's ring of chain tails holds at most 2^tail_log2 nodes (0: 24; a world of E entities
                            2^(log2 E + 14), at least 2^16); tail_exact 1: every world's ring holds 2^tail_log2 nodes (the
                            reference's --tail, src/rdma/rmain.cu:29-31, 1078-1085) */
    int32_t team_blocks; /* workgroups (of 256 threads) per world (0: batches of up to 8 worlds a workgroup per work item of
                            the largest world, at most 128 for a world alone or the GPU's share of each world (at least
                            32); larger batches a wave per work item, at most 16) */
    int32_t hosts_per_wave; /* hosts side by side on a wave's lanes (0: batches of up to 8 worlds the fewest of 2, 4, 8 or 16
                               whose items fit a wave each of the largest team; larger batches 64) */
    int32_t tail_exact;  /* see tail_log2 */
    int32_t serial;      /* 1: every switch slot runs its events in order on one lane, never the per-port parallel safe
                            path (the reference's --serial, src/rdma/rmain.cu:1090-1091, 1949); results are the same */
} trd_pack_cfg;

/* Worlds side by side on the GPU (the reference's --batch): one slot (the shortest link delay of
 * all), each world stepping its own slots. Replaces the device set-up of main
 * (src/rdma/rmain.cu:1292-1990). */
int trd_batch_create(trd_world *const *worlds, int32_t n, const trd_pack_cfg *cfg, trd_batch **out);

typedef struct trd_run_stats
{
    int64_t slots;        /* the most slots a world ran */
    int64_t sim_ns;       /* the latest world's end: its last slot's end (ns) */
    double wall_s;        /* the run's wall time */
    double kernel_s;      /* device time of the launches (HIP events) */
    uint32_t flows_done;
    int32_t launches;
} trd_run_stats;

/* Runs every world to its end: all flows done, the stop time (SIMULATOR_STOP_TIME), or
 * quiescence. Replaces main's launch loop (src/rdma/rmain.cu:2020-2170). */
int trd_batch_run(trd_batch *b, trd_run_stats *st);

/* a world's flows, by id: start and completion time minus start (-1: not finished) */
int trd_fct_get(const trd_batch *b, int32_t world, int64_t *start_ns, int64_t *fct_ns);
/* the counters of src/rdma/rmain.cu:2257-2258, by the reference's indices (40 entries) */
int trd_counters_get(const trd_batch *b, uint64_t *ev40);
/* The transmission trace (the reference's --txtrace=F --from=T0 --until=T1; src/rdma/rmain.cu:
 * 1064-1071, 1919-1927): every transmission start at a host or switch port at T0 <= t <= T1 (ns),
 * the first max_records of them (0: 64 Mi, the reference's). Call before trd_batch_run. */
int trd_txtrace_enable(trd_batch *b, int64_t from_ns, int64_t until_ns, uint64_t max_records);
/* Writes one world's trace after the run in the reference's format (src/rdma/rmain.cu:2262-2282):
 * "ns node dev kind src dst sport seq bytes ecn" per start (kind D A N P R; for PAUSE / RESUME src
 * is the queue and the rest 0), sorted by time, node, device. *records: lines written; *seen:
 * records the run made over all worlds (more than max_records: the rest were not kept). */
int trd_txtrace_write(const trd_batch *b, int32_t world, const char *path, uint64_t *records, uint64_t *seen);
/* the largest log2 of the batch's worlds' tail rings: what a run that lost tails (counter 9) had; a
 * rerun asks for more (the reference's --tail-auto, src/rdma/rmain.cu:2326-2363) */
int trd_batch_tail_log2(const trd_batch *b, int32_t *log2_nodes);
/* a world's slots run and device time (its team's constant-clock time, s) */
int trd_world_stats_get(const trd_batch *b, int32_t world, int64_t *slots, double *device_s);
void trd_batch_destroy(trd_batch *b);

/* The serving loop for RDMA requests (with include/twins.h's TW_REQ_* states): each request is a
 * world, packed into a device arena the server allocates once (no hipMalloc or hipFree while kernels
 * run); launches of the slot kernel run the resident, unfinished requests' worlds, small worlds (a
 * workgroup each, a quarter of the workgroups) and large ones (teams of up to 16) on two streams at
 * once; requests that
 * arrive during a launch (at most 1 s) join the next. Launches use at most max_blocks workgroups
 * (0: an eighth of the CUs; a TCP serving loop, tw_server_*, sized to leave them free runs side by
 * side). One host thread drives a server: submit, then poll until a request is done.
 * No counterpart in the reference (it has no serving loop). */
typedef struct trd_server trd_server;
typedef struct trd_server_cfg
{
    int32_t q3cap, tail_log2;  /* per request, as trd_pack_cfg (0: 4096 frames, at most 2^22 tail nodes) */
    int32_t hosts_per_wave;    /* 0: 16 */
    int32_t max_blocks;        /* workgroups the launches use at most, at least 2 (0: a CU in eight) */
    uint64_t arena_bytes;      /* 0: half of the free device memory at creation */
} trd_server_cfg;
typedef struct trd_server_stats
{
    int32_t waiting, resident, done, failed;
    uint64_t arena_used, arena_bytes, launches;
} trd_server_stats;
int trd_server_create(const trd_server_cfg *cfg, trd_server **out);
/* the world is copied; *id: the request's number (0, 1, ... in submission order) */
int trd_server_submit(trd_server *s, const trd_world *w, int32_t *id);
/* finished launches' results, admissions, the next launch; *st (may be null): the requests' states */
int trd_server_poll(trd_server *s, trd_server_stats *st);
int trd_server_state(const trd_server *s, int32_t id, int32_t *state);
/* a done request's flows (as trd_fct_get) and counters; its device memory returns to the arena. A
 * failed request's memory returns too, and the call returns an error. A request whose run met an
 * overwritten chain tail (counter 9) is run again by the server with larger tail rings (2^24, then
 * 2^26 nodes) before it is done. */
int trd_server_result(trd_server *s, int32_t id, int64_t *start_ns, int64_t *fct_ns, uint64_t *ev40);
void trd_server_destroy(trd_server *s);

const char *trd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
