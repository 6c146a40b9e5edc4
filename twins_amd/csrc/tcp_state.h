// Device-resident state of a batch of TCP worlds: the layout in HBM shared by the host-side packer
// (batch.cpp) and the gfx950 kernels (tcp_kernel.hip).
//
// A batch concatenates independent worlds. Every entity has a global id: hosts, switches,
// switch ports, links (one per sending queue: host h sends on link h, switch port p on link
// H + p, both world-local and offset by the world's link base), connections and messages. A
// frame carries world-local host ids and global connection ids.
//
// Layout choices (MI355X): per-connection state is an array of records split by owner (the
// sender's host and the receiver's host work on one connection in the same slot, each on its
// own record), each record aligned to 128-byte cache lines so one lane's field accesses stay in
// a few lines; queues, hosts, ports and link messages are arrays indexed by global id.
#pragma once
#include <cstdint>
#include "coherent.h"

// Device memory, for the device compile: pointers of the batch record carry the global address
// space, so the compiler emits global (not flat) loads and stores through them and what is derived
// from them (a flat access waits on LDS traffic too)
#if defined(__HIP_DEVICE_COMPILE__)
#define TW_G __attribute__((address_space(1)))
#else
#define TW_G
#endif

#if defined(__HIPCC__)
#define TW_HD __host__ __device__
#else
#define TW_HD
#endif

namespace tw
{

// Device-coherent accesses (coherent.h): what a team's workgroups hand each other from slot to slot
using twc::coh_get;
using twc::coh_load;
using twc::coh_put;
using twc::coh_store;

// A slot's work items on a team's waves: host items of `per` hosts side by side (one serial
// event loop per lane) and switch items of min(per, 64 / k) switches (a lane per port). The work
// items of a world with `per` hosts per item:
TW_HD inline int slot_items_at(int H, int S, int k, int per)
{
    const int smax = k > 0 && 64 / k > 0 ? 64 / k : 1, sp = per < smax ? per : smax;
    return (H + per - 1) / per + (S + sp - 1) / sp;
}
// the fewest hosts per item whose items do not outnumber the waves (a wave then runs one item per
// slot; fewer entities per item cost lanes but save the divergence of unlike event loops), at most
// 64 when even that does not fit
TW_HD inline void slot_items(int H, int S, int k, int waves, int *hpw, int *spw)
{
    int per = 1;
    while (per < 64 && slot_items_at(H, S, k, per) > waves)
        per++;
    const int smax = k > 0 && 64 / k > 0 ? 64 / k : 1;
    *hpw = per;
    *spw = per < smax ? per : smax;
}

constexpr int64_t SLOT_PS = 1000000, PROP_PS = 1000000, PS_PER_BYTE = 80;
constexpr uint32_t MSS = 1460;
constexpr int CTRL_FRAME = 42, SYN_FRAME = 46;
constexpr int MAX_TX = 300;    // frames one link can start in a slot (the model's cap)
constexpr int NRING = 3;       // link message ring: slots t-2, t-1 (read) and t (written)
// A link's message count word for slot t: the slot in the high 32 bits, then how many of its
// frames arrive before the end of slot t + 1, then how many it holds. The receiver takes the
// first ones in slot t + 1 and the rest in slot t + 2; frames start in order, one after another,
// so their arrivals increase and the first ones are a prefix.
TW_HD inline long long msg_word(int64_t t, int early, int sent)
{
    return (long long)((unsigned long long)t << 32) | (long long)early << 16 | sent;
}
TW_HD inline int64_t msg_slot(long long v) { return v >> 32; }
TW_HD inline int msg_early(long long v) { return (int)((v >> 16) & 0xFFFF); }
TW_HD inline int msg_sent(long long v) { return (int)(v & 0xFFFF); }
constexpr int MAX_ACTIVE = 256, MAX_DUE = 16;
constexpr int QDISC_CAP = 200, RED_K = 30;
constexpr int64_t MIN_RTO_PS = 1000000000, DELACK_PS = 10000000, CONN_TIMEOUT_PS = 10000000000LL,
                  CLOCK_GRAN_PS = 1000000, INIT_RTT_PS = 100000000;
constexpr uint32_t DELACK_COUNT = 2, INIT_CWND = 10 * MSS;
constexpr int DUPTHRESH = 3, DATA_RETRIES = 6;
constexpr double DCTCP_G = 0.0625;

constexpr uint8_t F_ACK = 1, F_SYN = 2, F_ECE = 4, F_ECT = 8, F_CE = 16, F_DATA = 32, F_CWR = 64, F_FIN = 128;
constexpr uint8_t IT_SACKED = 1, IT_LOST = 2, IT_RETX = 4;
constexpr uint8_t FIN_CLOSE = 1, FIN_SENT = 2, FIN_ACKED = 4, FIN_TW = 8;
enum : uint8_t { R_OPEN = 0, R_LAST_ACK, R_CLOSED };
enum : uint8_t { S_IDLE = 0, S_SYN_SENT, S_EST, S_DONE };
enum : uint8_t { C_OPEN = 0, C_DISORDER, C_CWR, C_RECOVERY, C_LOSS };
enum : uint8_t { E_IDLE = 0, E_CE_RCVD, E_SENDING_ECE, E_ECE_RCVD, E_CWR_SENT };
enum : uint8_t { CC_NEWRENO = 0, CC_DCTCP = 1, CC_CUBIC = 2 };
enum : int32_t { TOPO_FATTREE = 0, TOPO_LEAFSPINE = 1, TOPO_GRAPH = 2 };

// counters (the reference's ev[] numbering)
enum : int
{
    EV_DROP = 0,
    EV_MARK = 1,
    EV_RETX = 2,
    EV_RTO = 3,
    EV_TXCAP = 4,
    EV_ACTIVE_OVF = 5,
    EV_OOO_OVF = 6,
    EV_DUE_OVF = 7,
    EV_RETX_BY_STATE = 8, // 8..12
    EV_MAX_INFLIGHT = 13,
    EV_RTT_OVF = 14,
    // TW_PROF builds: host lanes' wall-clock ticks summed over hosts and slots: the slot's set-up
    // (inputs, timer scan), its event loop, of which the events' handling, and its end (active
    // list compaction); and the host events handled
    EV_PROF_H_SETUP = 15,
    EV_PROF_H_LOOP = 16,
    EV_PROF_H_EVENTS = 17,
    EV_PROF_H_END = 18,
    EV_RING_OVF = 19,
    EV_PROF_H_COUNT = 24,
    EV_UNDECIDED = 25,
    EV_PKT_EVENTS = 26, // frames started on a link (packet-events)
    // TW_PROF builds: wall-clock ticks summed over slots: slot period, slowest workgroup's host
    // work, slowest workgroup's switch work, slowest workgroup's total work
    EV_PROF_PERIOD = 27,
    EV_PROF_HOSTS = 28,
    EV_PROF_SWITCHES = 29,
    EV_PROF_WORK = 30,
    EV_N = 32
};

// A frame: 24 bytes of header then the event-order chain: RUNS runs, each a 16-bit code in the
// low half and a 16-bit count in the high half of a word, then the root key's two 16-bit words
// (high word first in memory): 64 bytes with 9 runs, 128 with 25.
template <int RUNS>
struct alignas(16) FrameT
{
    int64_t ts;       // enqueue time at a queue; arrival time at the next hop on a link
    uint32_t flow;    // connection (global id)
    uint32_t sa;      // data: first byte; otherwise the acknowledgement number
    uint16_t len;     // bytes on the wire
    uint16_t dst;     // destination host (world-local)
    uint8_t flags;
    uint8_t to_sender;
    uint8_t up;       // FatTree / leaf-spine: the ECMP index chosen at the sender
    uint8_t pad;
    uint32_t chain[RUNS + 1];
};
static_assert(sizeof(FrameT<9>) == 64 && sizeof(FrameT<25>) == 128, "frame layout");

// per connection, read-only
struct alignas(16) ConnStatic // (64 bytes: moved into LDS in 16-byte pieces with the event's record)
{
    int32_t src, dst; // global host ids
    uint32_t hash_data, hash_ack;
    int32_t m_first, m_cnt; // global message ids
    uint32_t it_cap, oo_cap;
    uint64_t it_off, oo_off; // ring offsets (sent list + RTT history; out-of-order blocks)
    int32_t world;
    uint8_t fin_on, pad[11];
};
static_assert(sizeof(ConnStatic) == 64, "ConnStatic: 64 bytes");

// the sending side (touched only by the sender's host)
struct alignas(128) SndState
{
    uint32_t tail, una, nxt, high, cwnd, ssthresh, dupacks, recover;
    uint32_t it_head, it_n, sack_b, lost_b, rtx_b, sacked_n, cand_n, cand;
    uint32_t hisack, h_head, h_n, cwnd_cnt, ack_rem, acked_total, acked_ecn, next_seq;
    int32_t m_due, m_wr, m_dn;
    uint8_t st, cong, fin, sack_seen, recover_active, first_partial, cwnd_limited, s_ce;
    uint8_t ecn_s, next_seq_flag, syn_tries, cub_found;
    int64_t srtt, rttvar, rto, rto_at, syn_at, timing_time, ecn_echo_seq, ecn_cwr_seq;
    double alpha;
    uint32_t cub_cnt, cub_lastmax, cub_origin, cub_endseq, cub_samples, cub_ackcnt, cub_tcpcwnd, pad0;
    double cub_k;
    int64_t cub_delaymin, cub_epoch, cub_roundstart, cub_lastack, cub_currrtt;
};

// the receiving side (touched only by the receiver's host)
struct alignas(128) RcvState
{
    uint32_t rcv_nxt, ooo_n, ooo_h, prior_rcv_nxt;
    uint8_t r_up, r_est, r_st, r_last_n, srv_syn_n, delack_count, ce_state, prior_flag;
    uint8_t reserved, ecn_r, pad[6];
    int64_t r_fin, r_synack_t, r_srtt, r_rttvar, r_last_at, r_frx_at, srv_syn_at, delack_at, ecn_ce_seq;
};

// MRG32k3a state (the layout of TwMrg in exact_math.h)
struct TwMrgState
{
    double s[6];
};

struct Due
{
    int64_t at;
    int32_t flow, kind; // 0 rto, 1 delack, 2 syn / deferred send, 3 SYN-ACK retx, 4 LastAck, 5 FIN+ACK retx
};

// one world of a batch
// One enqueue of the queue trace (the reference's --qtrace, src/model.h:312-320): the frame as it
// entered the queue, and the queue-disc length after it. Node, device and hosts follow on the host
// from the link and the connection.
struct QRec
{
    int64_t ts;
    int32_t link;   // global link id
    uint32_t flow;  // connection (global id)
    uint32_t sa;
    uint16_t len, qlen;
    uint8_t flags, to_sender, pad[6];
};
static_assert(sizeof(QRec) == 32, "trace record layout");
// the trace's buffer: enqueues at from <= now < until, the first cap of them (rec null: no trace)
struct QTrace
{
    QRec *rec;
    unsigned long long *cnt;
    uint64_t cap;
    int64_t from, until;
};

struct WorldDesc
{
    int32_t kind, k, hk, H, S, ecmp_n, leaves, spines, hpl, n_fail;
    int32_t host0, sw0, link0, L, conn0, nconn, msg0, M;
    int32_t rng0;  // first switch's MRG32k3a state (-1: per-flow ECMP)
    int32_t pad;
    int64_t max_slots;
    // topology files: offsets of the world's rows in the batch tables: host switch/port
    // [H], route offsets and counts [epoch][S][S] (offsets relative to the world's port
    // lists), port lists, down ports [epoch][S][k], failure times [n_fail]
    int64_t o_host, o_rt, o_rtp, o_pdn, o_fail;
};

// a world's progress (device-written)
struct WorldRun
{
    int64_t slot;         // the next slot to run
    int64_t slots_run;    // slots executed
    uint64_t next_wake;   // this slot: earliest host wake-up
    uint32_t act_cur, act_prev; // entities with activity: this slot, the previous one
    uint32_t flows_done;
    int32_t finished;
    int32_t error;
    int32_t team;               // workgroups of the team that ran it last
    unsigned long long ticks;   // device clock ticks (100 MHz) its team spent on it, over all launches
    unsigned long long ev[EV_N];
};

// a team of the launch plan: its size T, the hosts per wave work item it uses (0: per world), and
// the queue of worlds it takes from; first: its lowest workgroup. Its workgroups are those whose
// team_of entry is team << 8 | rank.
struct TeamPlan
{
    int32_t first, T, hpw, queue;
};

// a team's barrier and reductions (teams of several workgroups run one world), each word on its
// own cache line
struct alignas(128) TeamState
{
    unsigned arrive;
    unsigned pad0[31];
    unsigned gen;
    unsigned pad1[31];
    int32_t world, err;
    unsigned act;
    unsigned pad2;
    unsigned long long wake;
    unsigned long long prof[4]; // TW_PROF: this slot's maxima (hosts, switches, work); last ticket time
};

// The serving loop (tcp_server_kernel): requests are single-world batches resident in device
// memory. Queue q's requests are req[q * cap + i] in order; the host publishes tail[q] and the
// shutdown flag, and reads done[id] (1: finished, 2: failed), all in pinned host memory; the
// teams claim indices with next[q] (device memory).
struct ServerReq
{
    const void *batch; // BatchT in device memory
    int32_t id, pad;
};
struct ServerCtl
{
    ServerReq *req;
    unsigned *tail;
    int *shutdown;
    int *done;
    unsigned *next;
    int32_t nq, cap;
    unsigned long long idle_ticks; // a team waiting this long for a request gives its claim back
    int32_t *team_of;
    TeamPlan *plan;
    TeamState *teams;
    int64_t *scr_ts;
    uint32_t *scr_ref;
    uint8_t *scr_up;
};

// every device pointer of a batch (one copy in device memory, passed by pointer)
template <int RUNS>
struct BatchT
{
    typedef FrameT<RUNS> Frame;
    static constexpr int CW = RUNS + 1; // words of a chain
    int32_t nworlds, nhosts, nswitches, nlinks, nconns, nmsgs;
    WorldDesc TW_G *worlds;
    WorldRun TW_G *run;
    // connections
    ConnStatic TW_G *cs;
    SndState TW_G *snd;
    RcvState TW_G *rcv;
    uint32_t TW_G *rchain, *schain, *dchain, *fchain; // [conn][CW] timer chains
    uint32_t TW_G *it_start, *h_seq, *h_cnt, *oo_s, *oo_e;
    uint8_t TW_G *it_fl;
    int64_t TW_G *h_t;
    // messages
    int32_t TW_G *mg_conn;
    uint32_t TW_G *mg_end;
    int64_t TW_G *mg_start, *mg_est, *mg_done;
    int32_t TW_G *pend_msg;
    // hosts
    int32_t TW_G *host_world, *pend_next, *pend_end, *n_snd, *n_rcv, *act_snd, *act_rcv;
    uint8_t TW_G *host_cc;
    int64_t TW_G *wake;
    int64_t TW_G *tmin_s, *tmin_r; // per host: lower bounds of its connections' sender / receiver timers
    Due TW_G *due; // [host][MAX_DUE]
    // per link: the link whose frames reach the entity that sends on it (-1: an unused port)
    int32_t TW_G *in_link;
    int32_t TW_G *sw_world; // switch -> world
    // queues, one per link
    Frame TW_G *qring;
    uint64_t TW_G *q_base;
    int32_t TW_G *q_cap, *q_head, *q_size, *q_red_k, *q_limit;
    int64_t TW_G *q_link_free, *q_last_start;
    uint32_t TW_G *q_last_chain; // [link][CW]
    // link messages [NRING][link][MAX_TX] and their count words (msg_word)
    Frame TW_G *msg;
    long long TW_G *msg_cnt;
    // topology-file tables (all worlds)
    int32_t TW_G *g_host_sw, *g_host_port;
    uint32_t TW_G *g_rt_off;
    uint8_t TW_G *g_rt_n, *g_rt_ports, *g_port_dn;
    int64_t TW_G *g_fail_t;
    TwMrgState TW_G *sw_rng; // per-packet ECMP streams
    // kernel work queues: queue q holds worlds world_order[wq_off[q] .. wq_off[q] + wq_len[q] - 1] in
    // the order its teams take them, queue_next[q] counts those taken; the plan's teams (team_of:
    // workgroup -> team << 8 | rank in the team) and their barriers
    int32_t TW_G *world_order;
    int32_t TW_G *wq_off, *wq_len;
    int32_t TW_G *queue_next;
    int32_t nwq;
    int32_t TW_G *team_of;
    TeamPlan TW_G *plan;
    TeamState TW_G *teams;
    // switch-phase scratch, per wave: [wave][64 * 2 * MAX_TX] arrival times, refs, drawn routes
    int64_t TW_G *scr_ts;
    uint32_t TW_G *scr_ref;
    uint8_t TW_G *scr_up;
    int32_t scr_waves;
    QTrace qt; // the queue trace (off unless asked for)
};

} // namespace tw
