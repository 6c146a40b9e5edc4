// TwiNS-RDMA's packet model (the HPCC simulator's RDMA network: QbbNetDevice, RdmaHw with DCQCN,
// HPCC and DCTCP, SwitchNode and SwitchMmu with PFC and ECN), restated for MI355X. Every function
// here is a __host__ __device__ step of one entity in one slot; the slot kernel (rdma_kernel.hip)
// deals the entities of a world over the lanes of its workgroup, and a CPU build of the same code
// (tests/native/rdma_cpu) runs them one after the other. Reference: src/rdma/rdma.h and
// src/rdma/rstep.cuh; file:line citations name the function each step follows.
//
// The execution model is the reference's (src/rdma/rdma.h:12-16): slots of the shortest link delay D;
// a frame started in slot s arrives in slot s + 1 or later, filed by its link under its arrival slot,
// so within one slot every host and switch runs its own events (by time, simultaneous events by
// their event chains) from what earlier slots filed.
#pragma once
#include <cstdint>
#include <cstddef>
#include <climits>
#include <cmath>
#include "coherent.h"

// Device memory, for the device compile: pointers of the Sim record carry the global address space,
// so the compiler emits global (not flat) loads and stores through them and what is derived from
// them (a flat access waits on LDS traffic too)
#if defined(__HIP_DEVICE_COMPILE__)
#define RD_G __attribute__((address_space(1)))
#else
#define RD_G
#endif
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RD_HD __host__ __device__ __forceinline__
// (a cold path kept out of line: ties between simultaneous events are rare, and a chain comparison
// inlined at each of its call sites made the kernel's code far larger than the instruction cache)
#define RD_COLD __host__ __device__ __noinline__
#else
#define RD_HD inline
#define RD_COLD inline
#endif

namespace rd
{
// Device-coherent accesses (coherent.h). What a world's entities hand each other between slots
// goes only through these: link frames and their buckets' count words, the pending-slot flags,
// chain tail nodes, the lead-hash memo and the world's ticket. So a slot's team barrier needs no
// cache fences; each entity's own records stay in its workgroup's caches from slot to slot.
using twc::coh_get;
using twc::coh_load;
using twc::coh_put;
using twc::coh_store;
// ------------------------------------------------------------------ constants (src/rdma/rdma.h:101-106, 119-121)
constexpr int RUNS = 6;                 // chain runs kept inline
constexpr int NANC = 4;                 // ancestor lead hashes a chain carries (the memo's depths)
constexpr int ORD_BITS = 5;
constexpr uint32_t ORD_MAX = (1u << ORD_BITS) - 1;
constexpr uint64_t LEAD_MAX = (1ull << (32 - ORD_BITS)) - 2;
constexpr int MAXF = 256;               // frames one link delivers in one slot, at most
constexpr int MAX_HOP = 5;              // IntHeader::maxHop
constexpr int PG = 3;                   // the data priority group
constexpr uint32_t RESERVE = 4 * 1024, RESUME_OFFSET = 3 * 1024;
constexpr uint32_t PFC_SIZE = 9 + 20 + 14;
constexpr int TXT_N = 9216;             // frame sizes with a tabulated tx time
constexpr int ACKQ_CAP = 1024;          // a NIC's ACK queue
constexpr int MAX_DUE = 64;             // DCQCN timers one host runs in one slot
constexpr int NEV = 40;                 // counters

// counters (the reference's ev[] indices, printed in its summary line, src/rdma/rmain.cu:2257-2258)
enum : int
{
    EV_DROP = 0,
    EV_MARK = 1,
    EV_PAUSE = 2,
    EV_RESUME = 3,
    EV_TX_OVF = 4,
    EV_QUEUE_OVF = 5,
    EV_ACKQ_OVF = 6,
    EV_QPS_OVF = 7,
    EV_TIE = 8,
    EV_LOST = 9,
    EV_ACK_DONE = 10,
    EV_NACK = 11,
    EV_DUE_OVF = 12,
    EV_DATA_DONE = 13,
    EV_LATE = 32,
    EV_PKT = 38, // packet-events: frames sent (this build's; the reference's nout, src/rdma/rstep.cuh:1086, 2227)
};

enum : uint8_t
{
    K_DATA = 0,
    K_ACK,
    K_NACK,
    K_PAUSE,
    K_RESUME
};

// ------------------------------------------------------------------ records
// An event chain: run-length encoded leads back to the flow's StartApplication (src/rdma/rdma.h:108-131).
struct Run
{
    uint32_t code, cnt;
};
struct Chain
{
    Run r[RUNS];
    uint64_t tail; // older runs: a TailNode's ring position + 1 (0: none)
    uint64_t hash; // of the leads pushed since the root's StartApplication
    uint64_t anc[NANC];
    uint32_t root; // 1 + the flow's index in start order (0: none)
    uint32_t pad;
};
struct TailNode
{
    uint32_t code, cnt;
    uint64_t next, pos;
};

// A frame's contents; INT hops are IntHop words (lineRate:3 time:24 bytes:20 qlen:17 from the low bits).
struct Pkt
{
    uint64_t hop[MAX_HOP];
    uint32_t flow;
    uint32_t seq;
    uint16_t size;
    uint16_t nhop;
    uint8_t kind, ecn, cnp, q, inport, ord;
    uint16_t sport, src, dst;
};
static_assert(sizeof(Pkt) == 64, "Pkt");

struct LFrame
{
    int64_t ts; // arrival time
    Chain c;    // the chain of the event that started its transmission
    Pkt p;
};
// Frames and chains through coherent accesses, word by word as they lie in memory (the words built
// from the fields in registers: a record copied through memcpy would go to scratch memory).
static_assert(offsetof(Pkt, flow) == 40 && offsetof(Pkt, seq) == 44 && offsetof(Pkt, size) == 48 && offsetof(Pkt, nhop) == 50 &&
                  offsetof(Pkt, kind) == 52 && offsetof(Pkt, ecn) == 53 && offsetof(Pkt, cnp) == 54 && offsetof(Pkt, q) == 55 &&
                  offsetof(Pkt, inport) == 56 && offsetof(Pkt, ord) == 57 && offsetof(Pkt, sport) == 58 && offsetof(Pkt, src) == 60 &&
                  offsetof(Pkt, dst) == 62,
              "Pkt layout");
static_assert(offsetof(Chain, tail) == 8 * RUNS && offsetof(Chain, hash) == 8 * RUNS + 8 && offsetof(Chain, anc) == 8 * RUNS + 16 &&
                  offsetof(Chain, root) == 8 * RUNS + 16 + 8 * NANC && sizeof(Chain) == 8 * RUNS + 24 + 8 * NANC,
              "Chain layout");
RD_HD uint64_t pkt_w6(const Pkt &p)
{
    return (uint64_t)p.size | (uint64_t)p.nhop << 16 | (uint64_t)p.kind << 32 | (uint64_t)p.ecn << 40 | (uint64_t)p.cnp << 48 |
           (uint64_t)p.q << 56;
}
RD_HD uint64_t pkt_w7(const Pkt &p)
{
    return (uint64_t)p.inport | (uint64_t)p.ord << 8 | (uint64_t)p.sport << 16 | (uint64_t)p.src << 32 | (uint64_t)p.dst << 48;
}
template <typename PP>
RD_HD void pkt_put(PP d, const Pkt &p)
{
    auto w = reinterpret_cast<decltype(&d->hop[0])>(d);
    for (int i = 0; i < MAX_HOP; i++)
        coh_store(w + i, p.hop[i]);
    coh_store(w + 5, (uint64_t)p.flow | (uint64_t)p.seq << 32);
    coh_store(w + 6, pkt_w6(p));
    coh_store(w + 7, pkt_w7(p));
}
template <typename PP>
RD_HD Pkt pkt_get(PP f)
{
    Pkt p;
    auto w = reinterpret_cast<decltype(&f->hop[0])>(f);
    for (int i = 0; i < MAX_HOP; i++)
        p.hop[i] = coh_load(w + i);
    const uint64_t w5 = coh_load(w + 5), w6 = coh_load(w + 6), w7 = coh_load(w + 7);
    p.flow = (uint32_t)w5;
    p.seq = (uint32_t)(w5 >> 32);
    p.size = (uint16_t)w6;
    p.nhop = (uint16_t)(w6 >> 16);
    p.kind = (uint8_t)(w6 >> 32);
    p.ecn = (uint8_t)(w6 >> 40);
    p.cnp = (uint8_t)(w6 >> 48);
    p.q = (uint8_t)(w6 >> 56);
    p.inport = (uint8_t)w7;
    p.ord = (uint8_t)(w7 >> 8);
    p.sport = (uint16_t)(w7 >> 16);
    p.src = (uint16_t)(w7 >> 32);
    p.dst = (uint16_t)(w7 >> 48);
    return p;
}
static_assert(offsetof(Run, cnt) == 4 && sizeof(Run) == 8 && offsetof(Chain, pad) == offsetof(Chain, root) + 4, "Run, Chain words");
template <typename CP>
RD_HD void chain_put(CP d, const Chain &c)
{
    auto w = reinterpret_cast<decltype(&d->tail)>(d);
    for (int i = 0; i < RUNS; i++)
        coh_store(w + i, (uint64_t)c.r[i].code | (uint64_t)c.r[i].cnt << 32);
    coh_store(w + RUNS, c.tail);
    coh_store(w + RUNS + 1, c.hash);
    for (int i = 0; i < NANC; i++)
        coh_store(w + RUNS + 2 + i, c.anc[i]);
    coh_store(w + RUNS + 2 + NANC, (uint64_t)c.root);
}
template <typename CP>
RD_HD Chain chain_get(CP f)
{
    Chain c;
    auto w = reinterpret_cast<decltype(&f->tail)>(f);
    for (int i = 0; i < RUNS; i++)
    {
        const uint64_t x = coh_load(w + i);
        c.r[i].code = (uint32_t)x;
        c.r[i].cnt = (uint32_t)(x >> 32);
    }
    c.tail = coh_load(w + RUNS);
    c.hash = coh_load(w + RUNS + 1);
    for (int i = 0; i < NANC; i++)
        c.anc[i] = coh_load(w + RUNS + 2 + i);
    c.root = (uint32_t)coh_load(w + RUNS + 2 + NANC);
    c.pad = 0;
    return c;
}
// what a receiver looks at first: the arrival time, the size and the ordinal
struct FrameHead
{
    int64_t ts;
    uint32_t size;
    uint32_t ord;
};
RD_HD FrameHead frame_head(const LFrame RD_G *f)
{
    auto w = reinterpret_cast<const uint64_t RD_G *>(&f->p);
    return FrameHead{coh_load(&f->ts), (uint32_t)(uint16_t)coh_load(w + 6), (uint32_t)(uint8_t)(coh_load(w + 7) >> 8)};
}

struct Params
{
    int cc; // CC_MODE: 1 DCQCN, 3 HPCC, 8 DCTCP
    int int_size;
    uint32_t mtu, ack_interval, chunk;
    int64_t nack_interval;
    int var_win, rate_bound, ack_high_prio, ecn;
    double g, rate_first_cnp;
    int clamp;
    int64_t alpha_int, dec_int, rp_int;
    uint32_t rpg_threshold;
    uint64_t rai, rhai, min_rate, dctcp_rai;
    double u_target;
    uint32_t mi_thresh;
    int fast_react, sample_feedback;
    uint32_t int_multi;
    uint32_t buffer_size;
};

struct FlowInfo
{
    int src, dst; // global host ids
    uint64_t size;
    int64_t start;
    uint32_t sport, win;
    uint64_t base_rtt;
    Chain start_chain;
};

// a queue pair, sender side (RdmaQueuePair; src/rdma/rdma.h:174-229)
struct SndQP
{
    uint64_t size, snd_nxt, snd_una;
    uint64_t rate, max_rate;
    int64_t next_avail;
    uint32_t last_pkt, win;
    uint64_t base_rtt;
    uint8_t done, alpha_cnp, dec_cnp, first_cnp;
    uint32_t rp_stage;
    double alpha;
    uint64_t target;
    int64_t t_alpha, t_dec, t_rp; // pending DCQCN timers (-1: none)
    uint32_t last_update_seq;
    int32_t inc_stage;
    uint64_t cur_rate;              // HPCC
    uint32_t high_seq, ecn_cnt;     // DCTCP
    double u;                       // HPCC
    uint32_t batch, ca_state;       // DCTCP
    uint64_t hop[MAX_HOP];
    uint64_t tx_rate;               // the last tx time computed (qp_tx)
    uint32_t tx_size;
    int64_t tx_val;
};
struct QPChains
{
    Chain alpha, dec, rp;
};
struct RcvQP
{
    uint32_t expected, last_nack;
    int64_t nack_timer;
};

struct Route
{
    uint32_t off, n;
};

// an entity's share of the tail ring: positions next .. end - 1 reserved for it (taken from the
// shared counter TAIL_CHUNK at a time, so that a chain's spill waits for no global atomic)
struct TailAlloc
{
    uint64_t next, end;
    uint64_t wbits; // its world, in a position's top bits (see TailRing)
};
constexpr uint64_t TAIL_CHUNK = 16;
// Each world has a ring of tail nodes of its own (as a single run has), so that how fast a ring
// wraps does not depend on the worlds beside it: a position is world << TAIL_WSHIFT | (1 + the
// world's count), its node at tail[off + (count & mask)].
constexpr int TAIL_WSHIFT = 40;
struct TailRing
{
    uint64_t off, mask;
};

// a NIC (QbbNetDevice + RdmaEgressQueue state) and the host's flow starts
struct Host
{
    int64_t tc_at, ns_at; // pending TransmitComplete, m_nextSend (-1: none)
    uint32_t paused;
    int ackq_head, ackq_n;
    int qp_n, rr;
    int next_start, start_end;
    int64_t wake;    // its earliest event without input
    int64_t ob_last; // its out link's last bucket: slot << 16 | frames
    int64_t due_min; // at most its queue pairs' earliest DCQCN timer
    TailAlloc ta;
};

struct Port
{
    int64_t tc_at;
    uint32_t paused;
    int q0_head, q0_n, q3_head, q3_n;
    uint32_t bytes_total;
    uint64_t tx_bytes;
    uint32_t ing, hdrm, egr, mmu_paused; // SwitchMmu, this port as ingress and as egress
    int64_t ob_last;
    double rng[6]; // its ECN marking stream (MRG32k3a)
    TailAlloc ta;
};

struct Switch
{
    uint32_t shared_used, total_hdrm, total_rsrv;
    uint32_t max_shift; // the largest pfc shift of its ports (the lowest PFC threshold)
};

// a switch port's static set-up: its links, rate, headroom cap, PFC shift, ECN thresholds
struct PortCfg
{
    uint64_t bps;
    uint32_t hdrm_cap, shift, kmin, kmax;
    double pmax;
    int32_t link_in, link_out;
};

struct SwInfo
{
    uint32_t seed;  // its node id in its world (the ECMP hash seed)
    int hbase;      // its world's first host
    uint32_t rt_off;
    int pbase, deg; // its ports
};

// a directed link: its frames filed by arrival slot in nb buckets of cap frames
struct Link
{
    int64_t delay;
    uint64_t bps;
    uint64_t moff;
    uint32_t coff, nb, cap, rc;
};

// a sender's place in its out link's ring (src/rdma/rstep.cuh:45-66)
struct OutBucket
{
    int64_t lo, hi; // its slot's times (hi 0: none open)
    uint64_t base;
    uint32_t cnt, n;
    int link;
    int64_t last;
    uint32_t txs[2]; // its last two frame sizes and their tx times (a table lookup saved)
    int32_t txv[2];
    uint64_t moff;   // its link's ring (a copy: no load of the link per frame)
    uint32_t coff, nb, cap;
    int64_t delay;   // its link's delay
};

// a switch port's cursor over its slot's arrivals and its out bucket (per port, in HBM)
struct PortRun
{
    int64_t ats; // next arrival's time (INT64_MAX: none)
    uint32_t acode;
    int ai, an;
    int nout;
    uint64_t abase;
    OutBucket ob;
    int64_t idelay;   // its in-link's delay and rate class (read-only after the slot's set-up)
    uint32_t irc;
    uint64_t ibps;
    uint32_t isz[2];  // the in-link's tx times of full data frames and of ACKs (read-only likewise)
    int32_t itx[2];
    Pkt qhead;        // the head of its data queue when not empty (a copy: a dequeue waits for no load)
};

struct Due
{
    int64_t at;
    int flow, kind;
};

// a world: its entities and flows (global ids), its pending-slot ring, its stop
struct World
{
    int h0, nh, s0, ns, f0, nf;
    int pend_off, pend_n;
    int64_t max_slots;
};
// a world's run state
struct WorldRun
{
    int64_t slot, slots_run;
    uint64_t next_wake;
    uint32_t flows_done, act;
    int32_t finished, error;
    int64_t ticks;
};

// One transmission start of the trace (the reference's --txtrace, src/rdma/rdma.h:624-631): node
// (H + switch for switches) and device over all worlds; hosts of the frame (PAUSE / RESUME: the
// queue) as all-world ids. The writer turns them into one world's numbers.
struct TxRec
{
    int64_t ts;
    int32_t node, dev;
    uint32_t a, b, sport, seq;
    uint16_t size;
    uint8_t kind, ecn;
    uint32_t pad;
};
static_assert(sizeof(TxRec) == 40, "TxRec");
// the trace's buffer: starts at from <= now <= until, the first cap of them (rec null: no trace)
struct TxTrace
{
    TxRec *rec;
    unsigned long long *cnt;
    uint64_t cap;
    int64_t from, until;
};

struct Sim
{
    int H, S, F, W, L, R;
    int serial; // every switch slot on the serial path (--serial)
    int64_t D; // the slot (ns)
    const Params RD_G *Pw;
    const World RD_G *wd;
    WorldRun RD_G *run;
    const int RD_G *h_world, *h_link_in, *h_link_out;
    const SwInfo RD_G *si;
    const int RD_G *s_world;
    const PortCfg RD_G *pc;
    const Route RD_G *rt;
    const uint8_t RD_G *rtp;
    const FlowInfo RD_G *fi;
    const uint32_t RD_G *root_rank;
    const int RD_G *start_list;
    SndQP RD_G *qp;
    QPChains RD_G *qc;
    RcvQP RD_G *rx;
    Host RD_G *hs;
    Chain RD_G *h_tc, *h_ns;
    Pkt RD_G *ackq;
    int RD_G *qps;
    int qps_cap;
    Due RD_G *due; // [H][MAX_DUE] a host's timers of the slot
    Switch RD_G *sw;
    int RD_G *sw_busy; // [S] it ended its last slot busy (it runs the next one)
    Port RD_G *ps;
    Chain RD_G *p_tc;
    PortRun RD_G *pr;
    Pkt RD_G *q0, *q3;
    int q0_cap, q3_cap;
    const Link RD_G *li;
    LFrame RD_G *msg;
    uint32_t msg_bytes; // the link rings' size when below 2^31 (frames written 16 bytes a store), else 0
    int64_t RD_G *msg_cnt;
    const int32_t RD_G *txt; // [R][TXT_N]
    uint32_t RD_G *pend;
    TailNode RD_G *tail;
    unsigned long long RD_G *tail_cnt; // [W]
    const TailRing RD_G *rings;        // [W]
    unsigned long long RD_G *memo;
    uint64_t memo_mask;
    int64_t RD_G *fct;
    unsigned long long RD_G *ev;
    TxTrace txt_rec; // the transmission trace (off unless asked for)
};

// ------------------------------------------------------------------ atomics (device) / plain (CPU build)
RD_HD unsigned long long at_add64(unsigned long long *p, unsigned long long v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    unsigned long long o = *p;
    *p = o + v;
    return o;
#endif
}
RD_HD unsigned at_add32(unsigned *p, unsigned v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    unsigned o = *p;
    *p = o + v;
    return o;
#endif
}
RD_HD unsigned at_sub32(unsigned *p, unsigned v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicSub(p, v);
#else
    unsigned o = *p;
    *p = o - v;
    return o;
#endif
}
RD_HD unsigned long long at_cas64(unsigned long long *p, unsigned long long cmp, unsigned long long v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicCAS(p, cmp, v);
#else
    unsigned long long o = *p;
    if (o == cmp)
        *p = v;
    return o;
#endif
}
RD_HD void count(const Sim &s, int i) { at_add64(&s.ev[i], 1ull); }
RD_HD uint64_t mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// ------------------------------------------------------------------ ns-3.17 time (src/rdma/rstep.cuh:9-29)
// Seconds(x) in ns: int64x64_t(x) = floor(x) . floor(frac(x) 2^64), times 10^9, high word
RD_HD int64_t ns_seconds(double x)
{
    double hi = floor(x);
    uint64_t lo = (uint64_t)((x - hi) * 18446744073709551616.0);
    return (int64_t)hi * 1000000000ll + (int64_t)mulhi64(lo, 1000000000ull);
}
// DataRate(bps).CalculateTxTime(bytes) in ns
RD_HD int64_t tx_ns(uint64_t bps, uint32_t bytes) { return ns_seconds((double)bytes * 8 / (double)bps); }
RD_HD int64_t rc_tx(const Sim &s, uint32_t rc, uint64_t bps, uint32_t bytes)
{
    if (bytes < (uint32_t)TXT_N)
        return s.txt[(size_t)rc * TXT_N + bytes];
    return tx_ns(bps, bytes);
}
RD_HD int64_t link_tx(const Sim &s, int link, uint32_t bytes)
{
    const Link &l = s.li[link];
    if (bytes < (uint32_t)TXT_N)
        return s.txt[(size_t)l.rc * TXT_N + bytes];
    return tx_ns(l.bps, bytes);
}
// link_tx for a sender's out link, remembering its last two sizes (a sender sends few sizes)
RD_HD int64_t link_txo(const Sim &s, struct OutBucket &o, uint32_t bytes);
RD_HD uint64_t rate_of(double d) { return (uint64_t)d; }
RD_HD int ring_add(int a, int b, int n)
{
    int x = a + b;
    return x >= n ? x - n : x;
}

// ------------------------------------------------------------------ link rings (src/rdma/rstep.cuh:67-133; rdma.h:635-642)
// frames of `link` arriving in `slot`: their count, first one at *base
RD_HD int bucket(const Sim &s, int link, int64_t slot, uint64_t *base)
{
    const Link &l = s.li[link];
    int b = (int)(slot % (int64_t)l.nb);
    const int64_t v = coh_load(s.msg_cnt + l.coff + b);
    *base = l.moff + (uint64_t)b * l.cap;
    return (v >> 16) == slot ? (int)(v & 0xFFFF) : 0;
}
RD_HD void ob_open(const Sim &s, OutBucket &o, int link, int64_t last)
{
    const Link l = s.li[link];
    o.hi = 0;
    o.link = link;
    o.last = last;
    o.txs[0] = o.txs[1] = 0;
    o.moff = l.moff;
    o.coff = l.coff;
    o.nb = l.nb;
    o.cap = l.cap;
    o.delay = l.delay;
}
RD_HD int64_t link_txo(const Sim &s, OutBucket &o, uint32_t bytes)
{
    if (bytes == o.txs[0])
        return o.txv[0];
    if (bytes == o.txs[1])
        return o.txv[1];
    const int32_t t = (int32_t)link_tx(s, o.link, bytes);
    o.txs[1] = o.txs[0], o.txv[1] = o.txv[0];
    o.txs[0] = bytes, o.txv[0] = t;
    return t;
}
RD_HD void ob_flush(const Sim &s, OutBucket &o)
{
    if (o.hi == 0)
        return;
    o.last = (o.lo / s.D) << 16 | o.n;
    coh_store(s.msg_cnt + o.cnt, o.last);
    o.hi = 0;
}
// a place in the bucket of the frame's arrival slot (null: full, counted)
RD_HD LFrame RD_G *ob_push(const Sim &s, uint32_t *pend, int pend_n, int64_t ts, OutBucket &o)
{
    if (ts >= o.hi)
    {
        ob_flush(s, o);
        int64_t a = ts / s.D;
        int b = (int)(a % (int64_t)o.nb);
        o.cnt = o.coff + b;
        o.lo = a * s.D;
        o.hi = o.lo + s.D;
        o.base = o.moff + (uint64_t)b * o.cap;
        o.n = (o.last >> 16) == a ? (uint32_t)(o.last & 0xFFFF) : 0;
        coh_store(pend + a % pend_n, 1u);
    }
    if (o.n >= o.cap)
    {
        count(s, EV_TX_OVF);
        return nullptr;
    }
    return s.msg + o.base + o.n++;
}

// A frame's packet and chain, written coherently: 16 bytes a store (sc1 buffer stores) where the
// link rings fit a 32-bit offset, else 8-byte atomic stores; plain stores in the CPU build.
RD_HD void frame_put_pkt(const Sim &s, LFrame RD_G *fr, const Pkt &q)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (s.msg_bytes)
    {
        const __amdgpu_buffer_rsrc_t r = twc::coh_buffer((const void *)s.msg, s.msg_bytes);
        const uint32_t o = (uint32_t)((const char RD_G *)&fr->p - (const char RD_G *)s.msg);
        twc::coh_store16(r, o, q.hop[0], q.hop[1]);
        twc::coh_store16(r, o + 16, q.hop[2], q.hop[3]);
        twc::coh_store16(r, o + 32, q.hop[4], (uint64_t)q.flow | (uint64_t)q.seq << 32);
        twc::coh_store16(r, o + 48, pkt_w6(q), pkt_w7(q));
        return;
    }
#endif
    pkt_put(&fr->p, q);
}
RD_HD void frame_put_chain(const Sim &s, LFrame RD_G *fr, const Chain &c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (s.msg_bytes)
    {
        const __amdgpu_buffer_rsrc_t r = twc::coh_buffer((const void *)s.msg, s.msg_bytes);
        const uint32_t o = (uint32_t)((const char RD_G *)&fr->c - (const char RD_G *)s.msg);
        uint64_t w[RUNS + 2 + NANC + 1];
        for (int i = 0; i < RUNS; i++)
            w[i] = (uint64_t)c.r[i].code | (uint64_t)c.r[i].cnt << 32;
        w[RUNS] = c.tail;
        w[RUNS + 1] = c.hash;
        for (int i = 0; i < NANC; i++)
            w[RUNS + 2 + i] = c.anc[i];
        w[RUNS + 2 + NANC] = (uint64_t)c.root;
        constexpr int NW = RUNS + 2 + NANC + 1;
        for (int i = 0; i + 1 < NW; i += 2)
            twc::coh_store16(r, o + 8 * i, w[i], w[i + 1]);
        if (NW & 1)
            coh_store(reinterpret_cast<uint64_t RD_G *>(&fr->c) + NW - 1, w[NW - 1]);
        return;
    }
#endif
    chain_put(&fr->c, c);
}

// ------------------------------------------------------------------ event order (src/rdma/rstep.cuh:155-660)
RD_HD uint32_t ccode(int64_t lead, int ord)
{
    uint64_t l = (uint64_t)lead + 1;
    if (l > LEAD_MAX)
        l = LEAD_MAX;
    return (uint32_t)(l << ORD_BITS | (ORD_MAX - (ord < (int)ORD_MAX ? (uint32_t)ord : ORD_MAX)));
}
// splitmix64's finalizer over the lead (ordinals do not enter the hash)
RD_HD uint64_t chash(uint64_t h, uint32_t code)
{
    uint64_t z = h + 0x9e3779b97f4a7c15ull * ((code >> ORD_BITS) + 1ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// a tail position's node in the rings
RD_HD uint64_t tail_slot(const Sim &s, uint64_t pos)
{
    const TailRing r = s.rings[pos >> TAIL_WSHIFT];
    return r.off + (((pos & ((1ull << TAIL_WSHIFT) - 1)) - 1) & r.mask);
}
// the chain of an event scheduled with `code` by the event of chain par; the oldest run spills to
// its world's tail ring when the inline runs are full
RD_HD void cpush(const Sim &s, TailAlloc &ta, Chain *dst, uint32_t code, const Chain &par)
{
    Chain c;
    if (par.r[0].code == code)
    {
        c = par;
        if (c.r[0].cnt < 0xFFFFFFFFu)
            c.r[0].cnt++;
    }
    else
    {
        c.r[0].code = code;
        c.r[0].cnt = 1;
        for (int i = 1; i < RUNS; i++)
            c.r[i] = par.r[i - 1];
        c.root = par.root;
        c.tail = par.tail;
        c.pad = 0;
        if (par.r[RUNS - 1].code)
        {
            if (ta.next == ta.end)
            {
                ta.next = (at_add64(&s.tail_cnt[ta.wbits >> TAIL_WSHIFT], TAIL_CHUNK) + 1) | ta.wbits;
                ta.end = ta.next + TAIL_CHUNK;
            }
            const uint64_t pos = ta.next++;
            TailNode RD_G *n = s.tail + tail_slot(s, pos);
            coh_store(&n->code, par.r[RUNS - 1].code);
            coh_store(&n->cnt, par.r[RUNS - 1].cnt);
            coh_store(&n->next, par.tail);
            coh_store(&n->pos, pos);
            c.tail = pos;
        }
    }
    c.hash = chash(par.hash, code);
    c.anc[0] = par.hash;
    for (int i = 1; i < NANC; i++)
        c.anc[i] = par.anc[i - 1];
    *dst = c;
}

// Reading a chain's runs newest first, as if `code0` had been pushed on it (a candidate event's
// chain is compared before it exists): the pushed code (merged with the first run when equal), the
// inline runs, then the tail nodes, which may have been overwritten since (`lost`).
struct RunReader
{
    const Chain *c;
    int at;         // -1: the pushed code alone; 0 .. RUNS - 1: an inline run; RUNS + 1: in the tail
    uint64_t node;  // the tail node read (its position), 0: none yet
    uint64_t after; // the node after it
    uint32_t code;  // 0: the chain has ended (or its tail is lost)
    uint32_t left;  // repetitions left of the current run
    bool lost;
};
RD_HD void rr_fetch(const Sim &s, RunReader &r)
{
    if (r.at < RUNS)
    {
        const Run x = coh_get(&r.c->r[r.at]);
        r.code = x.code;
        r.left = x.cnt;
        if (x.code || r.at + 1 < RUNS)
            return; // an inline run, or the chain's end within them
        r.at = RUNS + 1; // every inline run is full: the tail follows
        r.node = coh_load(&r.c->tail);
    }
    if (!r.node)
    {
        r.code = 0;
        return;
    }
    const TailNode RD_G *tn = s.tail + tail_slot(s, r.node);
    const TailNode n = {coh_load(&tn->code), coh_load(&tn->cnt), coh_load(&tn->next), coh_load(&tn->pos)};
    if (n.pos != r.node)
    {
        r.lost = true; // the ring has wrapped over it
        r.code = 0;
        return;
    }
    r.code = n.code;
    r.left = n.cnt;
    r.after = n.next;
}
RD_HD void rr_open(const Sim &s, RunReader &r, const Chain &ch, uint32_t code0)
{
    r.c = &ch;
    r.node = r.after = 0;
    r.lost = false;
    if (!code0)
    {
        r.at = 0;
        rr_fetch(s, r);
        return;
    }
    const Run r0 = coh_get(&ch.r[0]);
    const bool joins = r0.code == code0;
    r.code = code0;
    r.at = joins ? 0 : -1;
    r.left = joins ? r0.cnt + 1 : 1;
}
RD_HD void rr_step(const Sim &s, RunReader &r)
{
    if (r.at > RUNS)
        r.node = r.after;
    else if (++r.at == RUNS)
    {
        r.at = RUNS + 1;
        r.node = coh_load(&r.c->tail);
    }
    rr_fetch(s, r);
}

// The memo of comparisons that leads alone decided, keyed by the two events' lead hashes: a
// small open-addressed table (four probes, a slot taken by compare-and-swap), shared by a world's
// entities, so that a comparison once walked is answered from the hashes (src/rdma/rstep.cuh:374-405
// describes the same table: its key and probe sequence are kept so that hits fall alike).
RD_HD uint64_t memo_key(uint64_t ha, uint64_t hb)
{
    const uint64_t k = chash(ha * 0x9e3779b97f4a7c15ull ^ hb, 0x5bd1e995u) & ~3ull;
    return k ? k : 4;
}
RD_HD int memo_find(const Sim &s, uint64_t ha, uint64_t hb)
{
    const uint64_t k = memo_key(ha, hb);
    for (uint64_t i = 0; i < 4; i++)
    {
        // (a plain read: a slot goes once from 0 to its key, so a stale 0 only costs a walk)
        const unsigned long long v = s.memo[(k + i) & s.memo_mask];
        if (!v)
            break; // (an empty slot ends the probe)
        if ((v & ~3ull) == k)
            return (v & 3) == 1 ? -1 : 1;
    }
    return 0;
}
RD_HD void memo_record(const Sim &s, uint64_t ha, uint64_t hb, int r)
{
    const uint64_t k = memo_key(ha, hb);
    for (uint64_t i = 0; i < 4; i++)
    {
        const unsigned long long was = at_cas64(&s.memo[(k + i) & s.memo_mask], 0ull, k | (r < 0 ? 1ull : 2ull));
        if (!was || (was & ~3ull) == k)
            return;
    }
}

// The order of two simultaneous events by their chains (< 0: a first, > 0: b, 0: undecided),
// walking both from the newest run (src/rdma/rstep.cuh:406-503 gives the rules): the first
// differing lead decides (the larger lead is the earlier ancestor); a chain that ends first runs
// first; with all leads equal the deepest differing ordinal decides, and distinct roots by their
// start order. *by_lead: decided by a lead or an end, i.e. by the leads alone; *inexact: a chain
// ended on an overwritten tail node.
RD_COLD int order_walk(const Sim &s, uint32_t pa, const Chain &a, uint32_t pb, const Chain &b, bool *inexact, bool *by_lead)
{
    RunReader A, B;
    rr_open(s, A, a, pa);
    rr_open(s, B, b, pb);
    int deep = 0;
    *inexact = *by_lead = false;
    while (A.code && B.code)
    {
        if ((A.code >> ORD_BITS) != (B.code >> ORD_BITS))
        {
            *by_lead = true;
            return A.code > B.code ? -1 : 1;
        }
        if (A.code != B.code)
            deep = (A.code & ORD_MAX) > (B.code & ORD_MAX) ? -1 : 1;
        if (A.at > RUNS && B.at > RUNS && A.node == B.node && A.left == B.left)
            return deep; // both at one place of one tail node: a single lineage from here
        const uint32_t m = A.left < B.left ? A.left : B.left;
        if (!(A.left -= m))
            rr_step(s, A);
        if (!(B.left -= m))
            rr_step(s, B);
    }
    const bool a_end = !A.code, b_end = !B.code;
    *inexact = (a_end && A.lost) || (b_end && B.lost);
    if (a_end && b_end)
    {
        const uint32_t a_root = coh_load(&a.root), b_root = coh_load(&b.root);
        return a_root && b_root && a_root != b_root ? (a_root < b_root ? -1 : 1) : deep;
    }
    if ((a_end ? A : B).lost)
        return deep;
    *by_lead = true;
    return a_end ? -1 : 1;
}
// the lead hashes of the event that code0 pushed on c would be, then of its ancestors
RD_HD void lead_hashes(uint64_t *h, uint32_t code0, const Chain &c)
{
    uint64_t anc[NANC];
    for (int i = 0; i < NANC; i++)
        anc[i] = coh_load(&c.anc[i]);
    const uint64_t ch = coh_load(&c.hash);
    h[0] = code0 ? chash(ch, code0) : ch;
    for (int i = 1; i <= NANC; i++)
        h[i] = !code0 ? anc[i - 1] : i == 1 ? ch : anc[i - 2];
}
// order_walk behind two shortcuts (src/rdma/rstep.cuh:504-575): lineages in lockstep (equal lead
// hashes, distinct roots) run in their roots' rank order; and the memo, probed at each depth
// while the leads agree. A walk decided by leads fills the memo at the depths it passed.
RD_COLD int order_events(const Sim &s, uint32_t pa, const Chain &a, uint32_t pb, const Chain &b, bool *inexact)
{
    uint64_t ha[NANC + 1], hb[NANC + 1];
    lead_hashes(ha, pa, a);
    lead_hashes(hb, pb, b);
    *inexact = false;
    const uint32_t a_root = coh_load(&a.root), b_root = coh_load(&b.root);
    if (a_root && b_root && a_root != b_root && ha[0] == hb[0])
        return s.root_rank[a_root - 1] < s.root_rank[b_root - 1] ? -1 : 1;
    RunReader A, B;
    rr_open(s, A, a, pa);
    rr_open(s, B, b, pb);
    int depth = 0; // the depths passed with equal leads
    for (;; depth++)
    {
        if (const int m = memo_find(s, ha[depth], hb[depth]))
            return m;
        if (depth == NANC || !A.code || !B.code || (A.code >> ORD_BITS) != (B.code >> ORD_BITS))
            break;
        if (!--A.left)
            rr_step(s, A);
        if (!--B.left)
            rr_step(s, B);
    }
    bool by_lead;
    const int r = order_walk(s, pa, a, pb, b, inexact, &by_lead);
    if (by_lead && !*inexact)
        for (int j = 0; j <= depth && j <= NANC; j++)
            memo_record(s, ha[j], hb[j], r);
    return r;
}

// A quick test on the chains' first runs (src/rdma/rstep.cuh:577-624): the newest HEAD runs of the
// chain code0 would make, compared by leads; 0 when they do not decide
constexpr int HEAD = 3;
RD_HD void head_runs(Run *out, uint32_t code0, const Chain &c)
{
    Run h[HEAD];
    for (int i = 0; i < HEAD; i++)
        h[i] = coh_get(&c.r[i]);
    const bool joins = code0 && h[0].code == code0;
    const int shift = code0 && !joins ? 1 : 0;
    for (int i = shift; i < HEAD; i++)
        out[i] = h[i - shift];
    if (code0)
    {
        out[0].code = code0;
        out[0].cnt = joins ? h[0].cnt + 1 : 1;
    }
}
RD_HD int heads_cmp(const Run *a, const Run *b)
{
    int i = 0, j = 0;
    uint32_t la = a[0].cnt, lb = b[0].cnt;
    while (i < HEAD && j < HEAD)
    {
        const uint32_t ca = a[i].code, cb = b[j].code;
        if (!ca || !cb)
            break;
        if ((ca >> ORD_BITS) != (cb >> ORD_BITS))
            return ca > cb ? -1 : 1;
        const uint32_t m = la < lb ? la : lb;
        la -= m;
        lb -= m;
        if (!la && ++i < HEAD)
            la = a[i].cnt;
        if (!lb && ++j < HEAD)
            lb = b[j].cnt;
    }
    return 0;
}
// does (ta, the event code pa pushes on a) run before (tb, the one pb pushes on b)? Times first,
// then the chains; an undecided tie keeps b (src/rdma/rstep.cuh:635-657)
// does (ta, the event code pa pushes on a) run before (tb, the one pb pushes on b)? Times first,
// then the chains; an undecided tie keeps b (src/rdma/rstep.cuh:635-657)
RD_COLD bool before(const Sim &s, int64_t ta, uint32_t pa, const Chain &a, int64_t tb, uint32_t pb, const Chain &b)
{
    if (ta != tb)
        return ta < tb;
    bool inexact;
    const int r = order_events(s, pa, a, pb, b, &inexact);
    if (inexact)
        count(s, EV_LOST);
    if (!r)
        count(s, EV_TIE);
    return r < 0;
}
// at one time: the root rule on the events' own lead hashes, then the head runs, then before()
// (the quick test src/rdma/rstep.cuh:626-633 applies to arrivals, TransmitCompletes and sends)
RD_COLD bool before_q(const Sim &s, int64_t t, uint32_t pa, const Chain &a, uint32_t pb, const Chain &b)
{
    const uint64_t a_hash = coh_load(&a.hash), b_hash = coh_load(&b.hash);
    const uint32_t a_root = coh_load(&a.root), b_root = coh_load(&b.root);
    const uint64_t ha = pa ? chash(a_hash, pa) : a_hash, hb = pb ? chash(b_hash, pb) : b_hash;
    int r;
    if (a_root && b_root && a_root != b_root && ha == hb)
        r = s.root_rank[a_root - 1] < s.root_rank[b_root - 1] ? -1 : 1;
    else
    {
        Run xa[HEAD], xb[HEAD];
        head_runs(xa, pa, a);
        head_runs(xb, pb, b);
        r = heads_cmp(xa, xb);
    }
    return r ? r < 0 : before(s, t, pa, a, t, pb, b);
}

// ------------------------------------------------------------------ frames (src/rdma/rstep.cuh:855-913)
RD_HD uint32_t node_ip(int id) { return 0x0b000001u + (id / 256) * 0x00010000u + (id % 256) * 0x00000100u; }
RD_HD uint32_t data_hdr(const Params &P) { return 48 + P.int_size; }
RD_HD uint32_t ack_size(const Params &P)
{
    int base = 12 + P.int_size, pad = 60 - 14 - 20 - base;
    return (pad > 0 ? pad : 0) + base + 20 + 14;
}
RD_HD uint64_t hop_time(uint64_t h) { return (h >> 3) & 0xFFFFFF; }
RD_HD uint64_t hop_bytes(uint64_t h) { return (h >> 27) & 0xFFFFF; }
RD_HD uint32_t hop_qlen(uint64_t h) { return (uint32_t)(h >> 47); }
RD_HD uint64_t hop_rate(uint64_t h)
{
    switch (h & 7)
    {
    case 0: return 25000000000ull;
    case 1: return 50000000000ull;
    case 2: return 100000000000ull;
    case 3: return 200000000000ull;
    case 4: return 400000000000ull;
    default: return 0;
    }
}
RD_HD uint64_t hop_make(uint64_t time, uint64_t bytes, uint32_t qlen, uint64_t rate, uint32_t multi)
{
    uint64_t lr = rate == 25000000000ull ? 0 : rate == 50000000000ull ? 1 : rate == 100000000000ull ? 2 : rate == 200000000000ull ? 3 : 4;
    uint64_t b = multi == 1 ? bytes >> 7 : bytes / (128ull * multi);
    uint64_t q = multi == 1 ? qlen / 80u : qlen / (80u * multi);
    return lr | (time & 0xFFFFFF) << 3 | (b & 0xFFFFF) << 27 | (q & 0x1FFFF) << 47;
}
RD_HD uint64_t hop_bytes_delta(uint64_t a, uint64_t b, uint32_t multi)
{
    uint64_t x = hop_bytes(a), y = hop_bytes(b);
    return (x >= y ? x - y : x + (1u << 20) - y) * 128 * multi;
}
RD_HD uint64_t hop_time_delta(uint64_t a, uint64_t b)
{
    uint64_t x = hop_time(a), y = hop_time(b);
    return x >= y ? x - y : x + (1u << 24) - y;
}
// SwitchNode::EcmpHash: MurmurHash3 x86_32 of three 32-bit words
RD_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
RD_HD uint32_t ecmp_hash(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t seed)
{
    const uint32_t w[3] = {w0, w1, w2};
    uint32_t h = seed;
    for (int i = 0; i < 3; i++)
    {
        uint32_t k = w[i] * 0xcc9e2d51u;
        k = rotl32(k, 15) * 0x1b873593u;
        h = rotl32(h ^ k, 13) * 5u + 0xe6546b64u;
    }
    h ^= 12u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    return h ^ (h >> 16);
}

// MRG32k3a (RngStream::RandU01): every intermediate an integer below 2^53, so contraction-safe
RD_HD double rand_u01(double *x)
{
    const double m1 = 4294967087.0, m2 = 4294944443.0, norm = 1.0 / (4294967087.0 + 1.0);
    double p1 = 1403580.0 * x[1] - 810728.0 * x[0];
    int32_t k = (int32_t)(p1 / m1);
    p1 -= k * m1;
    if (p1 < 0.0)
        p1 += m1;
    x[0] = x[1];
    x[1] = x[2];
    x[2] = p1;
    double p2 = 527612.0 * x[5] - 1370589.0 * x[3];
    k = (int32_t)(p2 / m2);
    p2 -= k * m2;
    if (p2 < 0.0)
        p2 += m2;
    x[3] = x[4];
    x[4] = x[5];
    x[5] = p2;
    return (p1 > p2) ? (p1 - p2) * norm : (p1 - p2 + m1) * norm;
}

// ------------------------------------------------------------------ host: the NIC
// A host's slot works on a NicSlot: the host record in registers, its out link's open bucket and
// the chain of the event that runs. The NIC's behaviour is the HPCC simulator's QbbNetDevice +
// RdmaHw (src/rdma/rstep.cuh:986-1647 restate it); here it is organised by concern: the egress
// pump (what the NIC sends next and when), the receive side (ACK / NACK replies), the queue pairs'
// lifecycle, and one rate controller per CC mode, each a small set of functions over the QP.
struct NicSlot
{
    const Sim *s;
    const Params *P;
    uint32_t *pend;
    int pend_n;
    int h;
    int64_t now, slot;
    Chain cur; // the running event's chain
    int ord;   // events it has scheduled
    int nout;  // frames sent this slot
    Host st;
    int link;  // out link
    int64_t ldelay;
    OutBucket ob;
};

// the chain of an event the running one schedules `lead` ns ahead (each takes the next ordinal)
RD_HD void nic_schedule(NicSlot &n, Chain *dst, int64_t lead) { cpush(*n.s, n.st.ta, dst, ccode(lead, n.ord++), n.cur); }

RD_HD uint64_t qp_unsent(const SndQP &q) { return q.size >= q.snd_nxt ? q.size - q.snd_nxt : 0; }
RD_HD bool qp_finished(const SndQP &q) { return q.snd_una >= q.size; }
// a frame's serialization time at `rate`, the last (rate, size) pair's kept with the QP
RD_HD int64_t qp_serial(SndQP &q, uint64_t rate, uint32_t bytes)
{
    if (q.tx_rate != rate || q.tx_size != bytes)
    {
        q.tx_val = tx_ns(rate, bytes);
        q.tx_rate = rate;
        q.tx_size = bytes;
    }
    return q.tx_val;
}
// the window: bytes in flight may not reach it (VAR_WIN scales it with the rate)
RD_HD bool qp_window_full(const Params &P, const SndQP &q)
{
    if (!q.win)
        return false;
    uint64_t lim = q.win;
    if (P.var_win)
    {
        lim = (uint64_t)q.win * q.rate / q.max_rate; // (q.win * rate stays below 2^64)
        lim = lim ? lim : 1;
    }
    return q.snd_nxt - q.snd_una >= lim;
}

// a transmission-trace record (src/rdma/rstep.cuh:915-942): out of line, off the path when tracing is off
// (the packet by value, as its words 6 and 7 and its sequence number: a packet passed by reference
// to this out-of-line function would be kept in scratch memory on the transmit path)
RD_COLD void trace_tx(const Sim &s, int64_t now, int node, int dev, uint32_t seq, uint64_t w6, uint64_t w7)
{
    const TxTrace &t = s.txt_rec;
    if (now < t.from || now > t.until)
        return;
    const unsigned long long i = at_add64(t.cnt, 1ull);
    if (i >= t.cap)
        return;
    const uint8_t kind = (uint8_t)(w6 >> 32), q = (uint8_t)(w6 >> 56);
    const uint16_t sport = (uint16_t)(w7 >> 16), src = (uint16_t)(w7 >> 32), dst = (uint16_t)(w7 >> 48);
    TxRec r;
    r.ts = now;
    r.node = node;
    r.dev = dev;
    r.kind = kind;
    r.size = (uint16_t)w6;
    r.ecn = (uint8_t)(w6 >> 40);
    r.pad = 0;
    const bool ctrl = kind >= K_PAUSE, fwd = kind == K_DATA;
    r.a = ctrl ? q : fwd ? src : dst;
    r.b = ctrl ? 0 : fwd ? dst : src;
    r.sport = ctrl ? 0 : sport;
    r.seq = ctrl ? 0 : seq;
    t.rec[i] = r;
}

// a frame onto the out link (TransmitStart): filed in its arrival bucket, the NIC busy until its
// last bit is out (the frame is written before the chain work, so a packet in a ring is copied
// straight into it)
RD_HD void nic_emit(NicSlot &n, const Pkt &p)
{
    const Sim &s = *n.s;
    const int64_t busy = link_txo(s, n.ob, p.size), lands = n.now + busy + n.ldelay;
    n.st.tc_at = n.now + busy;
    const int done_ord = n.ord; // the TransmitComplete's ordinal; the arrival's is the next
    n.ord += 2;
    n.nout++;
    LFrame RD_G *fr = ob_push(s, n.pend, n.pend_n, lands, n.ob);
    if (fr)
    {
        Pkt q = p;
        q.ord = (uint8_t)(done_ord < 254 ? done_ord + 1 : 255);
        coh_store(&fr->ts, lands);
        frame_put_pkt(s, fr, q);
        if (s.txt_rec.rec)
            trace_tx(s, n.now, n.h, 1, q.seq, pkt_w6(q), pkt_w7(q));
    }
    cpush(s, n.st.ta, &s.h_tc[n.h], ccode(busy, done_ord), n.cur);
    if (fr)
        frame_put_chain(s, fr, n.cur);
}

// The egress pump (DequeueAndTransmit + RdmaEgressQueue): when the link is free, the ACK queue
// first (unless its class is paused), else the queue pairs round-robin from the last one served.
// A QP is eligible with bytes left, its window open, the data class not paused and its pacing
// time reached. Finished QPs seen on the way leave the list (ns-3's lazy removal, done in the same
// pass that picks).
enum
{
    PICK_ACK = -1,
    PICK_NONE = -1024
};
// removes the finished QPs from list[from ..) (from: the first finished one); returns the new
// place of list index `keep` (PICK_NONE stays)
RD_HD int qp_list_prune(NicSlot &n, int *list, int from, int keep)
{
    const Sim &s = *n.s;
    int w = from;
    for (int i = from + 1; i < n.st.qp_n; i++)
    {
        if (qp_finished(s.qp[list[i]]))
            continue;
        if (i == keep)
            keep = w;
        list[w++] = list[i];
    }
    n.st.qp_n = w;
    return keep;
}
RD_HD int nic_pick(NicSlot &n)
{
    const Sim &s = *n.s;
    const int ack_class = n.P->ack_high_prio ? 0 : PG;
    if (n.st.ackq_n > 0 && !(n.st.paused >> ack_class & 1))
        return PICK_ACK;
    int *list = s.qps + (size_t)n.h * s.qps_cap;
    const int cnt = n.st.qp_n;
    const bool data_paused = n.st.paused >> PG & 1;
    int pick = PICK_NONE, first_done = -1;
    for (int k = 1, i = cnt ? n.st.rr % cnt : 0; k <= cnt; k++)
    {
        i = i + 1 < cnt ? i + 1 : 0;
        const SndQP &q = s.qp[list[i]];
        const bool ready = !data_paused && qp_unsent(q) > 0 && !qp_window_full(*n.P, q);
        if (!ready)
        {
            if (qp_finished(q) && (first_done < 0 || i < first_done))
                first_done = i;
            continue;
        }
        if (q.next_avail <= n.now)
        {
            pick = i;
            break;
        }
    }
    return first_done >= 0 ? qp_list_prune(n, list, first_done, pick) : pick;
}
// the next data segment of QP f (GetNxtPacket), then its pacing (PktSent, UpdateNextAvail)
RD_HD void nic_send_segment(NicSlot &n, int f)
{
    const Sim &s = *n.s;
    const Params &P = *n.P;
    SndQP &q = s.qp[f];
    const FlowInfo &fi = s.fi[f];
    const uint64_t left = qp_unsent(q);
    const uint32_t body = left < P.mtu ? (uint32_t)left : P.mtu;
    Pkt p = {};
    p.kind = K_DATA;
    p.flow = (uint32_t)f;
    p.src = (uint16_t)fi.src;
    p.dst = (uint16_t)fi.dst;
    p.sport = (uint16_t)fi.sport;
    p.seq = (uint32_t)q.snd_nxt;
    p.size = (uint16_t)(body + data_hdr(P));
    q.snd_nxt += body;
    nic_emit(n, p);
    q.last_pkt = p.size;
    q.next_avail = n.now + qp_serial(q, P.rate_bound ? q.rate : q.max_rate, p.size);
}
// nothing to send now: m_nextSend at the earliest pacing time of a QP with bytes left
RD_HD void nic_arm_next_send(NicSlot &n)
{
    const Sim &s = *n.s;
    const int *list = s.qps + (size_t)n.h * s.qps_cap;
    int64_t at = INT64_MAX;
    for (int i = 0; i < n.st.qp_n; i++)
    {
        const SndQP &q = s.qp[list[i]];
        if (qp_unsent(q) > 0 && q.next_avail < at)
            at = q.next_avail;
    }
    if (n.st.ns_at < 0 && at != INT64_MAX && at > n.now)
    {
        n.st.ns_at = at;
        nic_schedule(n, &s.h_ns[n.h], at - n.now);
    }
}
RD_HD void nic_pump(NicSlot &n)
{
    if (n.st.tc_at >= 0)
        return; // a transmission is in progress
    const int pick = nic_pick(n);
    const Sim &s = *n.s;
    if (pick == PICK_ACK)
    {
        const Pkt &a = s.ackq[(size_t)n.h * ACKQ_CAP + n.st.ackq_head]; // (the slot outlives the copy into the frame)
        n.st.ackq_head = ring_add(n.st.ackq_head, 1, ACKQ_CAP);
        n.st.ackq_n--;
        nic_emit(n, a);
    }
    else if (pick != PICK_NONE)
    {
        n.st.rr = pick;
        nic_send_segment(n, s.qps[(size_t)n.h * s.qps_cap + pick]);
    }
    else
        nic_arm_next_send(n);
}

// a QP's sending rate changes (ChangeRate): its pacing time moves by the difference of the last
// frame's serialization at the two rates, and an m_nextSend set later than it comes forward
RD_HD void qp_set_rate(NicSlot &n, SndQP &q, uint64_t rate)
{
    const int64_t was = qp_serial(q, q.rate, q.last_pkt), becomes = qp_serial(q, rate, q.last_pkt);
    q.next_avail += becomes - was;
    if (n.st.ns_at >= 0 && q.next_avail < n.st.ns_at)
    {
        const int64_t at = q.next_avail > n.now ? q.next_avail : n.now;
        n.st.ns_at = at;
        nic_schedule(n, &n.s->h_ns[n.h], at - n.now);
    }
    q.rate = rate;
}

// ---- HPCC (UpdateRateHp, one rate for the path): the path's utilization from the INT hops of
// the ACK against those of the last update, then a multiplicative or additive step
// the largest normalized utilization over the hops the ACK reports, and its hop's time delta; the
// QP's stored hops become the ACK's. false: no hop took part (all skipped by sample feedback).
RD_HD bool hpcc_path_load(const Params &P, SndQP &q, const Pkt &p, bool fast, double *U, uint64_t *tau_max)
{
    bool any = false;
    *U = 0;
    *tau_max = 0;
    // (MAX_HOP trips with static indices: the frame's hops stay in registers)
    for (int i = 0; i < MAX_HOP; i++)
    {
        if (i >= p.nhop)
            break;
        const uint64_t now_h = p.hop[i], was_h = q.hop[i];
        if (fast && P.sample_feedback && hop_qlen(now_h) == 0)
            continue;
        any = true;
        const uint64_t tau = hop_time_delta(now_h, was_h);
        const double bits = (double)(hop_bytes_delta(now_h, was_h, P.int_multi) * 8);
        const double link = (double)hop_rate(now_h);
        const uint32_t qn = hop_qlen(now_h) * 80 * P.int_multi, qw = hop_qlen(was_h) * 80 * P.int_multi;
        const double util = bits / ((double)tau * 1e-9) / link + (double)(qn < qw ? qn : qw) * (double)q.max_rate / link / (double)q.win;
        if (util > *U)
        {
            *U = util;
            *tau_max = tau;
        }
        q.hop[i] = now_h;
    }
    return any;
}
RD_HD void hpcc_react(NicSlot &n, SndQP &q, double U, uint64_t tau, bool fast)
{
    const Params &P = *n.P;
    if (tau > q.base_rtt)
        tau = q.base_rtt;
    // the EWMA of U over a base RTT (the fork's build contracts u (B - tau) + U tau into one fma)
    q.u = fma(q.u, (double)(q.base_rtt - tau), U * (double)tau) / (double)q.base_rtt;
    const double over = q.u / P.u_target;
    const bool mult = over >= 1 || (uint32_t)q.inc_stage >= P.mi_thresh;
    uint64_t rate = mult ? rate_of((double)q.cur_rate / over) + P.rai : q.cur_rate + P.rai;
    rate = rate < P.min_rate ? P.min_rate : rate > q.max_rate ? q.max_rate : rate;
    qp_set_rate(n, q, rate);
    if (!fast)
    {
        q.cur_rate = rate;
        q.inc_stage = mult ? 0 : q.inc_stage + 1;
    }
}
RD_HD void hpcc_on_ack(NicSlot &n, SndQP &q, const Pkt &p, bool fast)
{
    const uint32_t sent_to = (uint32_t)q.snd_nxt;
    if (!q.last_update_seq)
    {
        // the first ACK only records the path
        q.last_update_seq = sent_to;
        for (int i = 0; i < MAX_HOP; i++)
            if (i < p.nhop)
                q.hop[i] = p.hop[i];
        return;
    }
    double U;
    uint64_t tau;
    if (p.nhop <= MAX_HOP && hpcc_path_load(*n.P, q, p, fast, &U, &tau))
        hpcc_react(n, q, U, tau, fast);
    if (!fast && sent_to > q.last_update_seq)
        q.last_update_seq = sent_to;
}

// ---- DCQCN (Mellanox): a CNP arms the alpha and rate-decrease timers on the QP's first one; the
// three timers re-arm themselves (their deadlines feed the host's earliest-timer bound)
RD_HD void dcqcn_arm(NicSlot &n, int64_t *deadline, Chain *chain, int64_t lead)
{
    *deadline = n.now + lead;
    if (*deadline < n.st.due_min)
        n.st.due_min = *deadline;
    nic_schedule(n, chain, lead);
}
RD_HD uint64_t dcqcn_cut(const Params &P, const SndQP &q)
{
    const uint64_t r = rate_of((double)q.rate * (1 - q.alpha / 2));
    return r > P.min_rate ? r : P.min_rate;
}
RD_HD void dcqcn_on_cnp(NicSlot &n, int f, SndQP &q)
{
    const Params &P = *n.P;
    q.alpha_cnp = q.dec_cnp = 1;
    if (!q.first_cnp)
        return;
    q.first_cnp = 0;
    q.alpha = 1;
    q.alpha_cnp = 0;
    dcqcn_arm(n, &q.t_alpha, &n.s->qc[f].alpha, P.alpha_int);
    dcqcn_arm(n, &q.t_dec, &n.s->qc[f].dec, P.dec_int + 1);
    q.rate = rate_of((double)q.rate * P.rate_first_cnp);
    q.target = q.rate;
}
RD_HD void dcqcn_alpha_tick(NicSlot &n, int f, SndQP &q)
{
    const Params &P = *n.P;
    q.alpha = q.alpha_cnp ? fma(1 - P.g, q.alpha, P.g) : (1 - P.g) * q.alpha;
    q.alpha_cnp = 0;
    dcqcn_arm(n, &q.t_alpha, &n.s->qc[f].alpha, P.alpha_int);
}
RD_HD void dcqcn_decrease_tick(NicSlot &n, int f, SndQP &q)
{
    const Params &P = *n.P;
    dcqcn_arm(n, &q.t_dec, &n.s->qc[f].dec, P.dec_int);
    if (!q.dec_cnp)
        return;
    if (q.rp_stage || P.clamp)
        q.target = q.rate;
    q.rate = dcqcn_cut(P, q);
    q.rp_stage = 0;
    q.dec_cnp = 0;
    dcqcn_arm(n, &q.t_rp, &n.s->qc[f].rp, P.rp_int);
}
RD_HD void dcqcn_increase_tick(NicSlot &n, int f, SndQP &q)
{
    const Params &P = *n.P;
    dcqcn_arm(n, &q.t_rp, &n.s->qc[f].rp, P.rp_int);
    if (q.rp_stage >= P.rpg_threshold)
    {
        const uint64_t t = q.target + (q.rp_stage == P.rpg_threshold ? P.rai : P.rhai);
        q.target = t > q.max_rate ? q.max_rate : t;
    }
    q.rate = rate_of((double)q.rate / 2) + rate_of((double)q.target / 2);
    q.rp_stage++;
}

// ---- DCTCP over RDMA (HandleAckDctcp): alpha per window of segments, one cut per window with a
// mark, an additive step per window without
// closes the window of ACKs when one passes its end: alpha from its marked share; true if so
RD_HD bool dctcp_window(const Params &P, SndQP &q, const Pkt &p)
{
    if (p.seq <= q.last_update_seq)
        return false;
    if (q.last_update_seq)
    {
        const double marked = (double)q.ecn_cnt / (double)q.batch;
        q.alpha = fma(1 - P.g, q.alpha, P.g * (marked < 1.0 ? marked : 1.0));
        q.ecn_cnt = 0;
        q.batch = (uint32_t)((q.snd_nxt - (uint64_t)p.seq) / P.mtu + 1);
    }
    else
        q.batch = (uint32_t)(q.snd_nxt / P.mtu + 1);
    q.last_update_seq = (uint32_t)q.snd_nxt;
    return true;
}
RD_HD void dctcp_on_ack(const Params &P, SndQP &q, const Pkt &p)
{
    if (p.cnp)
        q.ecn_cnt++;
    const bool window_end = dctcp_window(P, q, p);
    if (q.ca_state == 1 && p.seq > q.high_seq)
        q.ca_state = 0; // the reduced window is acknowledged
    if (q.ca_state == 0 && p.cnp)
    {
        q.rate = dcqcn_cut(P, q); // (the same rate cut: rate (1 - alpha/2), at least the minimum)
        q.ca_state = 1;
        q.high_seq = (uint32_t)q.snd_nxt;
    }
    if (q.ca_state == 0 && window_end)
    {
        const uint64_t r = q.rate + P.dctcp_rai;
        q.rate = r > q.max_rate ? q.max_rate : r;
    }
}

// ---- the queue pairs' lifecycle
// an ACK or NACK (ReceiveAck): the QP's acknowledged point moves; the last byte completes it (its
// flow completion time is recorded, its timers end); a NACK rewinds sending to it (go-back-N);
// then the CC mode's reaction and the pump
RD_HD void nic_on_ack(NicSlot &n, const Pkt &p, unsigned *flows_done)
{
    const Sim &s = *n.s;
    const Params &P = *n.P;
    const int f = (int)p.flow;
    SndQP &q = s.qp[f];
    if (q.done)
    {
        count(s, EV_ACK_DONE);
        return;
    }
    q.snd_una = p.seq > q.snd_una ? p.seq : q.snd_una;
    const bool completes = qp_finished(q);
    if (completes)
    {
        q.done = 1;
        q.t_alpha = q.t_dec = q.t_rp = -1;
        s.fct[f] = n.now - s.fi[f].start;
        at_add32(flows_done, 1u);
    }
    if (p.kind == K_NACK)
        q.snd_nxt = q.snd_una;
    switch (P.cc)
    {
    case 1:
        if (p.cnp && !completes)
            dcqcn_on_cnp(n, f, q);
        break;
    case 3:
        if (p.seq > q.last_update_seq || P.fast_react)
            hpcc_on_ack(n, q, p, p.seq <= q.last_update_seq);
        break;
    case 8:
        if (!completes)
            dctcp_on_ack(P, q, p);
        break;
    }
    nic_pump(n);
}
// a data frame at the receiver (ReceiveUdp, ReceiverCheckSeq): in order, it advances the
// expected byte and is acknowledged at the ACK interval or a chunk's end; ahead of it, a NACK
// unless one for the same point is still within its interval; old data is ignored
enum
{
    RX_ACK,
    RX_NACK,
    RX_QUIET
};
RD_HD int rx_check(const Params &P, RcvQP &r, const Pkt &p, int64_t now)
{
    const uint32_t body = p.size - data_hdr(P);
    if (p.seq == r.expected)
    {
        r.expected += body;
        return r.expected >= P.ack_interval || (P.chunk && r.expected % P.chunk == 0) ? RX_ACK : RX_QUIET;
    }
    if (p.seq < r.expected || (now < r.nack_timer && r.last_nack == r.expected))
        return RX_QUIET;
    r.nack_timer = now + P.nack_interval;
    r.last_nack = r.expected;
    return RX_NACK;
}
RD_HD void nic_on_data(NicSlot &n, const Pkt &p)
{
    const Sim &s = *n.s;
    RcvQP &r = s.rx[p.flow];
    if (r.expected >= s.fi[p.flow].size)
        count(s, EV_DATA_DONE);
    const int v = rx_check(*n.P, r, p, n.now);
    if (v == RX_QUIET)
        return;
    if (v == RX_NACK)
        count(s, EV_NACK);
    if (n.st.ackq_n >= ACKQ_CAP)
    {
        count(s, EV_ACKQ_OVF);
        return;
    }
    // the reply carries the data frame's addresses and hops, the expected byte and the CE mark as CNP
    Pkt &a = s.ackq[(size_t)n.h * ACKQ_CAP + ring_add(n.st.ackq_head, n.st.ackq_n, ACKQ_CAP)];
    a = p;
    a.kind = v == RX_ACK ? K_ACK : K_NACK;
    a.seq = r.expected;
    a.cnp = p.ecn ? 1 : 0;
    a.ecn = 0;
    a.size = (uint16_t)ack_size(*n.P);
    n.st.ackq_n++;
    nic_pump(n);
}
// a flow starts (StartApplication, AddQueuePair): a fresh QP at line rate joins the host's list
RD_HD void nic_add_qp(NicSlot &n, int f)
{
    const Sim &s = *n.s;
    const FlowInfo &fi = s.fi[f];
    const uint64_t line = s.li[n.link].bps;
    SndQP q = {};
    q.size = fi.size;
    q.win = fi.win;
    q.base_rtt = fi.base_rtt;
    q.rate = q.max_rate = q.target = line;
    q.alpha = 1;
    q.first_cnp = 1;
    q.t_alpha = q.t_dec = q.t_rp = -1;
    q.next_avail = n.now;
    if (n.P->cc == 3)
    {
        q.cur_rate = line;
        q.u = 1;
    }
    s.qp[f] = q;
    if (n.st.qp_n >= s.qps_cap)
    {
        count(s, EV_QPS_OVF);
        return;
    }
    s.qps[(size_t)n.h * s.qps_cap + n.st.qp_n++] = f;
    nic_pump(n);
}

RD_HD const Chain &due_chain(const Sim &s, const Due &d)
{
    const QPChains &qc = s.qc[d.flow];
    return d.kind == 0 ? qc.alpha : d.kind == 1 ? qc.dec : qc.rp;
}
RD_HD int64_t due_now(const Sim &s, const Due &d)
{
    const SndQP &q = s.qp[d.flow];
    return d.kind == 0 ? q.t_alpha : d.kind == 1 ? q.t_dec : q.t_rp;
}
RD_HD int64_t qp_timer_min(const Sim &s, int h, int n)
{
    const int *list = s.qps + (size_t)h * s.qps_cap;
    int64_t mn = INT64_MAX;
    for (int i = 0; i < n; i++)
    {
        const SndQP &q = s.qp[list[i]];
        const int64_t t[3] = {q.t_alpha, q.t_dec, q.t_rp};
        for (int k = 0; k < 3; k++)
            mn = t[k] >= 0 && t[k] < mn ? t[k] : mn;
    }
    return mn;
}

// One host's slot (src/rdma/rstep.cuh:1649-2112 for a host with one NIC): its events in ns-3's
// order - the candidates are its next arrival, its TransmitComplete, its m_nextSend, its next DCQCN
// timer and its next flow start; the earliest runs, simultaneous ones by their chains. Returns
// whether the host is active (sent, or holds frames); *wake: its earliest event without input.
RD_HD bool host_step(const Sim &s, int h, int64_t slot, unsigned *flows_done, uint32_t *pend, int pend_n, int64_t *wake)
{
    const int64_t D = s.D, t0 = slot * D, t1 = t0 + D;
    const Params &P = s.Pw[s.h_world[h]];
    const int lin = s.h_link_in[h];
    uint64_t abase;
    const int an = bucket(s, lin, slot, &abase);
    if (an == 0 && s.hs[h].wake >= t1)
    {
        *wake = s.hs[h].wake;
        return false;
    }
    NicSlot c;
    c.s = &s;
    c.P = &P;
    c.pend = pend;
    c.pend_n = pend_n;
    c.h = h;
    c.slot = slot;
    c.now = t0;
    c.nout = 0;
    c.ord = 0;
    c.st = s.hs[h];
    c.link = s.h_link_out[h];
    c.ldelay = s.li[c.link].delay;
    ob_open(s, c.ob, c.link, c.st.ob_last);
    const int64_t in_delay = s.li[lin].delay;
    uint32_t in_sz = 0; // the in-link's last frame size and its tx time
    int64_t in_tx = 0;
    // the DCQCN timers due in this slot, in event order (their intervals are at least a slot: none
    // set in it falls due in it)
    Due *due = s.due + (size_t)h * MAX_DUE;
    int ndue = 0;
    if (P.cc == 1 && c.st.due_min < t1)
    {
        const int *list = s.qps + (size_t)h * s.qps_cap;
        int64_t mn = INT64_MAX;
        for (int i = 0; i < c.st.qp_n; i++)
        {
            const SndQP &q = s.qp[list[i]];
            const int64_t t[3] = {q.t_alpha, q.t_dec, q.t_rp};
            for (int k = 0; k < 3; k++)
                mn = t[k] >= 0 && t[k] < mn ? t[k] : mn;
            for (int k = 0; k < 3; k++)
                if (t[k] >= 0 && t[k] < t1)
                {
                    if (ndue == MAX_DUE)
                    {
                        count(s, EV_DUE_OVF);
                        continue;
                    }
                    Due x = {t[k], list[i], k};
                    int j = ndue++;
                    while (j > 0 && before(s, x.at, 0, due_chain(s, x), due[j - 1].at, 0, due_chain(s, due[j - 1])))
                    {
                        due[j] = due[j - 1];
                        j--;
                    }
                    due[j] = x;
                }
        }
        c.st.due_min = mn;
    }
    int di = 0, ai = 0;
    int sf = -1, sf_at = -1;
    int64_t sf_t = INT64_MAX;
    // the next arrival's frame (written by the link's sender in an earlier slot: coherent reads):
    // its time and packet read once, together (its chain when it runs)
    const LFrame RD_G *af = nullptr;
    FrameHead ah = {INT64_MAX, 0, 0};
    Pkt apk;
    int a_at = -1;
    for (;;)
    {
        if (a_at != ai)
        {
            a_at = ai;
            af = ai < an ? s.msg + abase + ai : nullptr;
            if (af)
            {
                apk = pkt_get(&af->p);
                ah = FrameHead{coh_load(&af->ts), apk.size, apk.ord};
            }
            else
                ah = FrameHead{INT64_MAX, 0, 0};
        }
        if (af && ah.size != in_sz)
        {
            in_sz = ah.size;
            in_tx = link_tx(s, lin, in_sz);
        }
        const uint32_t acode = af ? ccode(in_tx + in_delay, ah.ord) : 0;
        while (di < ndue && due_now(s, due[di]) != due[di].at)
            di++; // cancelled or moved
        if (sf_at != c.st.next_start)
        {
            sf_at = c.st.next_start;
            sf = sf_at < c.st.start_end ? s.start_list[sf_at] : -1;
            sf_t = sf >= 0 ? s.fi[sf].start : INT64_MAX;
        }
        // candidate times: 0 arrival, 1 TransmitComplete, 2 m_nextSend, 3 DCQCN timer, 4 flow start
        int64_t tt[5];
        tt[0] = ah.ts;
        tt[1] = c.st.tc_at >= 0 && c.st.tc_at < t1 ? c.st.tc_at : INT64_MAX;
        tt[2] = c.st.ns_at >= 0 && c.st.ns_at < t1 ? c.st.ns_at : INT64_MAX;
        tt[3] = di < ndue ? due[di].at : INT64_MAX;
        tt[4] = sf_t < t1 ? sf_t : INT64_MAX;
        // only the candidates at the earliest time are compared (in candidate order)
        int64_t bt = INT64_MAX;
        for (int k = 0; k < 5; k++)
            bt = tt[k] < bt ? tt[k] : bt;
        int bk = -1;
        for (int k = 0; k < 5 && bt != INT64_MAX; k++)
        {
            const int64_t t = tt[k];
            if (t != bt)
                continue;
            if (bk < 0)
            {
                bk = k;
                continue;
            }
            // a tie: the quick prefixes where both are arrivals, TransmitCompletes or sends, else the chains
            const uint32_t pk = k == 0 ? acode : 0, pb = bk == 0 ? acode : 0;
            const Chain &ck = k == 0 ? af->c : k == 1 ? s.h_tc[h] : k == 2 ? s.h_ns[h] : k == 3 ? due_chain(s, due[di]) : s.fi[sf].start_chain;
            const Chain &cb = bk == 0 ? af->c : bk == 1 ? s.h_tc[h] : bk == 2 ? s.h_ns[h] : bk == 3 ? due_chain(s, due[di]) : s.fi[sf].start_chain;
            const bool b4 = (k <= 2 && bk <= 2) ? before_q(s, t, pk, ck, pb, cb) : before(s, t, pk, ck, bt, pb, cb);
            if (b4)
                bk = k;
        }
        if (bk < 0)
            break;
        if (bt < t0)
            count(s, EV_LATE);
        c.now = bt;
        if (bk == 0)
            cpush(s, c.st.ta, &c.cur, acode, chain_get(&af->c));
        else
            c.cur = bk == 1 ? s.h_tc[h] : bk == 2 ? s.h_ns[h] : bk == 3 ? due_chain(s, due[di]) : s.fi[sf].start_chain;
        c.ord = 0;
        if (bk == 0)
        {
            const Pkt p = apk;
            ai++;
            if (p.kind == K_DATA)
                nic_on_data(c, p);
            else if (p.kind == K_ACK || p.kind == K_NACK)
                nic_on_ack(c, p, flows_done);
            else if (p.kind == K_PAUSE)
                c.st.paused |= 1u << p.q;
            else
            {
                c.st.paused &= ~(1u << p.q);
                nic_pump(c);
            }
        }
        else if (bk == 1)
        {
            c.st.tc_at = -1;
            nic_pump(c);
        }
        else if (bk == 2)
        {
            c.st.ns_at = -1;
            nic_pump(c);
        }
        else if (bk == 3)
        {
            // a DCQCN timer: 0 alpha update, 1 rate-decrease check, 2 rate increase
            const Due x = due[di++];
            SndQP &q = s.qp[x.flow];
            if (x.kind == 0)
                dcqcn_alpha_tick(c, x.flow, q);
            else if (x.kind == 1)
                dcqcn_decrease_tick(c, x.flow, q);
            else
                dcqcn_increase_tick(c, x.flow, q);
        }
        else
        {
            c.st.next_start++;
            nic_add_qp(c, sf);
        }
    }
    bool act = c.nout > 0;
    int64_t w = INT64_MAX;
    if (c.st.next_start < c.st.start_end)
        w = s.fi[s.start_list[c.st.next_start]].start;
    ob_flush(s, c.ob);
    c.st.ob_last = c.ob.last;
    if (c.st.ackq_n > 0 || c.st.tc_at >= 0)
        act = true;
    if (c.st.ns_at >= 0 && c.st.ns_at < w)
        w = c.st.ns_at;
    if (P.cc == 1)
    {
        if (c.st.due_min < t1)
            c.st.due_min = qp_timer_min(s, h, c.st.qp_n);
        if (c.st.due_min < w)
            w = c.st.due_min;
    }
    if (c.st.tc_at >= 0 && c.st.tc_at < w)
        w = c.st.tc_at;
    c.st.wake = w;
    s.hs[h] = c.st;
    if (c.nout)
        at_add64(&s.ev[EV_PKT], (unsigned long long)c.nout);
    *wake = w;
    return act;
}

// ------------------------------------------------------------------ switch: ports and the MMU
// A switch's slot works on a SwSlot: its ports (in device memory, or LDS in a safe slot), the
// ports' per-slot state, the switch-wide MMU counters in registers and the running event's chain.
// The behaviour is the HPCC simulator's SwitchNode + SwitchMmu (src/rdma/rstep.cuh:2117-2467
// restate it); here it is organised as the MMU's accounting (admission, release, PFC decisions),
// the egress port's queues and transmitter, and the ingress path that ties them together.
struct SwSlot
{
    const Sim *s;
    const Params *P;
    uint32_t *pend;
    int pend_n;
    int x, node;
    Port *pt;      // its ports (in HBM)
    PortRun *w;
    Chain *tc;     // the pending TransmitCompletes' chains
    const PortCfg *pc;
    Switch sw;
    int64_t now;
    Chain cur;     // made on first use: cpush(cur_code, *cur_src)
    const Chain *cur_src;
    uint32_t cur_code;
    bool cur_ok;
    int pbase; // its first port in the per-port arrays
    bool safe; // a safe slot: this context runs one egress port's events (see switch_slot)
    int ep;    // the running event's port (its tail positions: one writer per port and slot)
    int ord;
};
// the running event's chain, pushed the first time an event it schedules needs it
RD_HD const Chain &sw_event_chain(SwSlot &c)
{
    if (!c.cur_ok)
    {
        c.cur_ok = true;
        if (!c.cur_code)
            c.cur = *c.cur_src; // (a TransmitComplete's chain, the port's own)
        else
            cpush(*c.s, c.pt[c.ep].ta, &c.cur, c.cur_code, chain_get(c.cur_src)); // (an arrival's, in its frame)
    }
    return c.cur;
}

// ---- the MMU (shared buffer with per-port reserve, headroom and dynamic PFC thresholds)
// a port's use of the shared pool: its ingress bytes above the reserve
RD_HD uint32_t mmu_shared_of(const Port &p) { return p.ing > RESERVE ? p.ing - RESERVE : 0; }
// the PFC threshold of an ingress port: the free shared buffer, shifted by the port's factor
RD_HD uint32_t mmu_pfc_limit(const SwSlot &c, int port)
{
    const uint32_t free_shared = c.P->buffer_size - c.sw.total_hdrm - c.sw.total_rsrv - c.sw.shared_used;
    return free_shared >> c.pc[port].shift;
}
// a data frame entering through `inport` (CheckIngressAdmission, UpdateIngressAdmission): false
// if it is dropped; else its bytes go to the reserve, then the shared pool, then headroom
RD_HD bool mmu_admit(SwSlot &c, int inport, uint32_t size)
{
    Port &in = c.pt[inport];
    const uint32_t lim = mmu_pfc_limit(c, inport);
    if (size + in.hdrm > c.pc[inport].hdrm_cap && size + mmu_shared_of(in) > lim)
        return false;
    const uint32_t after = in.ing + size;
    if (after > RESERVE && after - RESERVE > lim)
    {
        in.hdrm += size;
        return true;
    }
    in.ing = after;
    if (after > RESERVE)
        c.sw.shared_used += size < after - RESERVE ? size : after - RESERVE;
    return true;
}
// a data frame leaving (RemoveFromIngressAdmission): headroom first, then the shared pool and
// the reserve
RD_HD void mmu_release(SwSlot &c, int inport, uint32_t size)
{
    Port &in = c.pt[inport];
    const uint32_t hd = in.hdrm < size ? in.hdrm : size, rest = size - hd, sh = mmu_shared_of(in);
    in.hdrm -= hd;
    in.ing -= rest;
    c.sw.shared_used -= rest < sh ? rest : sh;
}
// ECN at dequeue (ShouldSendCN): above Kmax always, between Kmin and Kmax with a probability
// rising to Pmax, drawn from the port's own stream
RD_HD bool mmu_ecn_mark(SwSlot &c, int port)
{
    const PortCfg &pc = c.pc[port];
    const uint32_t q = c.pt[port].egr;
    if (q > pc.kmax)
        return true;
    if (q <= pc.kmin)
        return false;
    return rand_u01(c.pt[port].rng) < pc.pmax * (double)(q - pc.kmin) / (double)(pc.kmax - pc.kmin);
}

// ---- an egress port: the control queue (PFC frames; ACKs under ack_high_prio), the data queue
// and the transmitter
// a frame onto the port's link (TransmitStart): as a NIC's, the frame first, then the chains
RD_HD void port_emit(SwSlot &c, int port, const Pkt &p)
{
    const Sim &s = *c.s;
    PortRun &r = c.w[port];
    const int64_t busy = link_txo(s, r.ob, p.size), lands = c.now + busy + r.ob.delay;
    c.pt[port].tc_at = c.now + busy;
    const int done_ord = c.ord;
    c.ord += 2;
    r.nout++;
    LFrame RD_G *fr = ob_push(s, c.pend, c.pend_n, lands, r.ob);
    if (fr)
    {
        Pkt q = p;
        q.ord = (uint8_t)(done_ord < 254 ? done_ord + 1 : 255);
        coh_store(&fr->ts, lands);
        frame_put_pkt(s, fr, q);
        if (s.txt_rec.rec)
            trace_tx(s, c.now, c.node, port + 1, q.seq, pkt_w6(q), pkt_w7(q));
    }
    const Chain &ev = sw_event_chain(c);
    cpush(s, c.pt[c.ep].ta, &c.tc[port], ccode(busy, done_ord), ev);
    if (fr)
        frame_put_chain(s, fr, ev);
}
// a frame into one of the port's queues (0: control, else data); the data queue's head is kept
// in the port's slot state
RD_HD void port_push(SwSlot &c, int port, int queue, const Pkt &p)
{
    const Sim &s = *c.s;
    Port &pt = c.pt[port];
    const size_t pi = (size_t)c.pbase + port;
    const bool ctrl = queue == 0;
    const int n = ctrl ? pt.q0_n : pt.q3_n, cap = ctrl ? s.q0_cap : s.q3_cap;
    if (n >= cap)
    {
        count(s, EV_QUEUE_OVF);
        return;
    }
    if (ctrl)
    {
        s.q0[pi * s.q0_cap + ring_add(pt.q0_head, n, s.q0_cap)] = p;
        pt.q0_n = n + 1;
    }
    else
    {
        s.q3[pi * s.q3_cap + ring_add(pt.q3_head, n, s.q3_cap)] = p;
        if (!n)
            c.w[port].qhead = p;
        pt.q3_n = n + 1;
    }
    pt.bytes_total += p.size;
}
// the control queue's head onto an idle link
RD_HD void port_send_ctrl(SwSlot &c, int port)
{
    const Sim &s = *c.s;
    Port &pt = c.pt[port];
    if (pt.tc_at >= 0)
        return;
    const Pkt &p = s.q0[((size_t)c.pbase + port) * s.q0_cap + pt.q0_head]; // (the slot outlives the copy into the frame)
    pt.q0_head = ring_add(pt.q0_head, 1, s.q0_cap);
    pt.q0_n--;
    pt.bytes_total -= p.size;
    pt.tx_bytes += p.size;
    port_emit(c, port, p);
}
// a PAUSE or RESUME for the data class out of `port` (SendPfc), ahead of everything queued there
RD_HD void port_pfc(SwSlot &c, int port, bool pause)
{
    count(*c.s, pause ? EV_PAUSE : EV_RESUME);
    Pkt f = {};
    f.kind = pause ? K_PAUSE : K_RESUME;
    f.q = PG;
    f.size = PFC_SIZE;
    port_push(c, port, 0, f);
    if (c.pt[port].q0_n > 0)
        port_send_ctrl(c, port);
}
// a data-queue frame leaves through `port`: the MMU lets go of it (in a safe slot only its ingress
// counter moves), it may be ECN-marked, its ingress port may RESUME, an HPCC frame gets this hop's
// INT record, then it is sent
RD_HD void port_send_data(SwSlot &c, int port, Pkt p)
{
    const Sim &s = *c.s;
    const Params &P = *c.P;
    Port &pt = c.pt[port];
    if (c.safe)
        at_sub32(&c.pt[p.inport].ing, (uint32_t)p.size); // (no headroom, no pause anywhere in a safe slot)
    else
        mmu_release(c, p.inport, p.size);
    pt.egr -= p.size;
    if (P.ecn && mmu_ecn_mark(c, port))
    {
        p.ecn = 3;
        count(s, EV_MARK);
    }
    Port &in = c.pt[p.inport];
    if (!c.safe && in.mmu_paused && in.hdrm == 0)
    {
        const uint32_t sh = mmu_shared_of(in);
        if (sh == 0 || sh + RESUME_OFFSET <= mmu_pfc_limit(c, p.inport))
        {
            port_pfc(c, p.inport, false);
            in.mmu_paused = 0;
        }
    }
    if (p.kind == K_DATA && P.cc == 3 && p.nhop < 0xFFFF)
    {
        const uint64_t rec = hop_make((uint64_t)c.now, pt.tx_bytes, pt.bytes_total, c.pc[port].bps, P.int_multi);
        const int slot_i = p.nhop % MAX_HOP;
        // (static indices only: a frame indexed at run time would leave the registers for scratch memory)
        for (int k = 0; k < MAX_HOP; k++)
            p.hop[k] = k == slot_i ? rec : p.hop[k];
        p.nhop++;
    }
    pt.tx_bytes += p.size;
    port_emit(c, port, p);
}
// the port's link is free (DequeueAndTransmit): control frames first, then the data queue unless
// the data class is paused
RD_HD void port_pump(SwSlot &c, int port)
{
    const Sim &s = *c.s;
    Port &pt = c.pt[port];
    if (pt.tc_at >= 0)
        return;
    if (pt.q0_n > 0)
    {
        port_send_ctrl(c, port);
        return;
    }
    if (pt.q3_n == 0 || (pt.paused >> PG & 1))
        return;
    PortRun &r = c.w[port];
    const Pkt p = r.qhead;
    pt.q3_head = ring_add(pt.q3_head, 1, s.q3_cap);
    pt.q3_n--;
    pt.bytes_total -= p.size;
    if (pt.q3_n > 0) // (the next head's load overlaps this frame's transmission)
        r.qhead = s.q3[((size_t)c.pbase + port) * s.q3_cap + pt.q3_head];
    port_send_data(c, port, p);
}

// ---- the ingress path
// the egress port of a data frame or an ACK (-1: no route) and its queue: equal-cost routes to
// the destination host, chosen by the 5-tuple's hash with the switch's seed
RD_HD int sw_route(const Sim &s, const Params &P, int x, const Pkt &p, int *q)
{
    const SwInfo &si = s.si[x];
    const bool fwd = p.kind == K_DATA;
    const int a = (fwd ? p.src : p.dst) - si.hbase, b = (fwd ? p.dst : p.src) - si.hbase;
    *q = !fwd && P.ack_high_prio ? 0 : PG;
    const Route rt = s.rt[si.rt_off + b];
    if (!rt.n)
        return -1;
    const uint32_t ports_word = fwd ? (p.sport | 100u << 16) : (100u | (uint32_t)p.sport << 16);
    return s.rtp[rt.off + ecmp_hash(node_ip(a), node_ip(b), ports_word, si.seed) % rt.n];
}
// a frame through `inport` bound for `out` (queue q): the MMU's admission (a drop, or a PAUSE when
// the ingress port crosses its limit), then cut-through to an idle port with nothing queued, or
// the queue and the pump
RD_HD void sw_ingress(SwSlot &c, int inport, Pkt p, int out, int q)
{
    if (out < 0)
        return;
    if (q && c.safe)
    {
        at_add32(&c.pt[inport].ing, (uint32_t)p.size); // (below every limit all slot: shared_used follows at its end)
        c.pt[out].egr += p.size;
    }
    else if (q)
    {
        if (!mmu_admit(c, inport, p.size))
        {
            count(*c.s, EV_DROP);
            return;
        }
        c.pt[out].egr += p.size;
        Port &in = c.pt[inport];
        if (!in.mmu_paused && (in.hdrm > 0 || mmu_shared_of(in) >= mmu_pfc_limit(c, inport)))
        {
            port_pfc(c, inport, true);
            in.mmu_paused = 1;
        }
    }
    p.inport = (uint8_t)inport;
    const Port &o = c.pt[out];
    const bool idle = o.tc_at < 0 && o.q0_n == 0 && o.q3_n == 0 && !(o.paused >> PG & 1);
    if (q && idle)
    {
        port_send_data(c, out, p);
        return;
    }
    port_push(c, out, q, p);
    port_pump(c, out);
}

// a frame's tx time on a port's in-link
RD_HD int64_t port_in_tx(const Sim &s, const PortRun &r, uint32_t bytes)
{
    return bytes == r.isz[0] ? r.itx[0] : bytes == r.isz[1] ? r.itx[1] : rc_tx(s, r.irc, r.ibps, bytes);
}
// a port's next arrival of the slot
RD_HD void sw_peek(const Sim &s, const SwSlot &c, int p)
{
    PortRun &w = c.w[p];
    w.ats = INT64_MAX;
    if (w.ai < w.an)
    {
        const FrameHead h = frame_head(s.msg + w.abase + w.ai);
        w.ats = h.ts;
        w.acode = ccode(port_in_tx(s, w, h.size) + w.idelay, h.ord);
    }
}

RD_HD void sw_ctx(SwSlot &c, const Sim &s, int x, int64_t slot, uint32_t *pend, int pend_n, Port *pt, PortRun *w)
{
    const SwInfo &si = s.si[x];
    c.s = &s;
    c.P = &s.Pw[s.s_world[x]];
    c.pend = pend;
    c.pend_n = pend_n;
    c.x = x;
    c.node = s.H + x;
    c.pt = pt;
    c.w = w;
    c.tc = s.p_tc + si.pbase;
    c.pc = s.pc + si.pbase;
    c.pbase = si.pbase;
    c.sw = s.sw[x];
    c.safe = false;
    c.cur_ok = false;
    c.ord = 0;
    c.ep = 0;
    c.now = slot * s.D;
}
// a port's slot set-up: its arrivals' cursor and its out bucket
RD_HD void sw_port_open(const Sim &s, SwSlot &c, int p, int64_t slot)
{
    PortRun &r = c.w[p];
    r.ai = 0;
    r.nout = 0;
    r.an = bucket(s, c.pc[p].link_in, slot, &r.abase);
    {
        const Link &li = s.li[c.pc[p].link_in];
        r.idelay = li.delay;
        r.irc = li.rc;
        r.ibps = li.bps;
        r.isz[0] = c.P->mtu + data_hdr(*c.P);
        r.isz[1] = ack_size(*c.P);
        r.itx[0] = (int32_t)rc_tx(s, r.irc, r.ibps, r.isz[0]);
        r.itx[1] = (int32_t)rc_tx(s, r.irc, r.ibps, r.isz[1]);
    }
    if (c.pt[p].q3_n > 0)
        r.qhead = s.q3[((size_t)c.pbase + p) * s.q3_cap + c.pt[p].q3_head];
    ob_open(s, r.ob, c.pc[p].link_out, c.pt[p].ob_last);
}
// a port's slot end: its bucket's count out; busy (sent, or holds frames)?
RD_HD bool sw_port_close(const Sim &s, SwSlot &c, int p, unsigned long long *nout)
{
    ob_flush(s, c.w[p].ob);
    c.pt[p].ob_last = c.w[p].ob.last;
    *nout += (unsigned long long)c.w[p].nout;
    return c.w[p].nout > 0 || c.pt[p].q0_n > 0 || c.pt[p].q3_n > 0 || c.pt[p].tc_at >= 0;
}

// All of a switch's events of the slot in order on one lane (the reference's serial path,
// src/rdma/rstep.cuh:3032-3189): candidates are each port's next arrival and its TransmitComplete,
// in port order, arrival first; ties by the chains; undecided ties keep the earlier candidate.
RD_HD void sw_serial_events(const Sim &s, SwSlot &c, int deg, int64_t t1)
{
    for (int p = 0; p < deg; p++)
        sw_peek(s, c, p);
    for (;;)
    {
        int best = -1, bk = 0;
        int64_t bt = INT64_MAX;
        for (int p = 0; p < deg; p++)
        {
            const int64_t ta = c.w[p].ats, tcat = c.pt[p].tc_at;
            for (int k = 0; k < 2; k++)
            {
                const int64_t t = k ? tcat : ta;
                if (k ? (tcat < 0 || tcat >= t1 || tcat > bt) : (ta == INT64_MAX || ta > bt))
                    continue;
                if (t < bt || best < 0)
                {
                    best = p, bk = k, bt = t;
                    continue;
                }
                const uint32_t ca = k ? 0 : c.w[p].acode, cb = bk ? 0 : c.w[best].acode;
                const Chain &xa = k ? c.tc[p] : s.msg[c.w[p].abase + c.w[p].ai].c;
                const Chain &xb = bk ? c.tc[best] : s.msg[c.w[best].abase + c.w[best].ai].c;
                if (before_q(s, t, ca, xa, cb, xb))
                    best = p, bk = k;
            }
        }
        if (best < 0)
            break;
        const int p = best;
        c.now = bt;
        c.ord = 0;
        c.ep = p;
        c.cur_ok = false;
        if (bk == 1)
        {
            c.cur_src = &c.tc[p];
            c.cur_code = 0;
            sw_event_chain(c); // (the TransmitComplete's chain is overwritten by the next transmission)
            c.pt[p].tc_at = -1;
            port_pump(c, p);
            continue;
        }
        const LFrame RD_G *f = s.msg + c.w[p].abase + c.w[p].ai;
        const Pkt pk = pkt_get(&f->p);
        c.cur_code = c.w[p].acode;
        c.cur_src = &f->c;
        c.w[p].ai++;
        sw_peek(s, c, p);
        if (pk.kind == K_PAUSE || pk.kind == K_RESUME)
        {
            if (pk.kind == K_PAUSE)
                c.pt[p].paused |= 1u << pk.q;
            else
            {
                c.pt[p].paused &= ~(1u << pk.q);
                port_pump(c, p);
            }
        }
        else
        {
            int q = PG;
            const int out = sw_route(s, *c.P, c.x, pk, &q);
            sw_ingress(c, p, pk, out, q);
        }
    }
}

// One switch's slot on one lane, its ports in HBM: the serial path for any number of ports.
// Returns whether the switch is busy.
RD_HD bool switch_step(const Sim &s, int x, int64_t slot, uint32_t *pend, int pend_n)
{
    const SwInfo &si = s.si[x];
    SwSlot c;
    sw_ctx(c, s, x, slot, pend, pend_n, s.ps + si.pbase, s.pr + si.pbase);
    for (int p = 0; p < si.deg; p++)
        sw_port_open(s, c, p, slot);
    sw_serial_events(s, c, si.deg, (slot + 1) * s.D);
    s.sw[x] = c.sw;
    bool act = false;
    unsigned long long nout = 0;
    for (int p = 0; p < si.deg; p++)
        act |= sw_port_close(s, c, p, &nout);
    if (nout)
        at_add64(&s.ev[EV_PKT], nout);
    return act;
}

// ------------------------------------------------------------------ a switch's slot by the lanes of a wave
// A switch of at most KS ports runs its slot from a workspace (LDS on the GPU): its ports, its
// arrivals' routes and times. The lanes load the ports and route the arrivals; then the slot is
// either safe (src/rdma/rstep.cuh:2760-3030) - no port paused by the MMU, none with headroom bytes,
// and every port below the lowest PFC threshold the slot can reach even if all its arrivals stay,
// so that no MMU decision (drop, headroom, PAUSE, RESUME) can occur and egress ports interact
// only through sums in their ingress counters - and each egress port's events (the arrivals routed
// to it, merged from the ingress ports in event order, and its TransmitCompletes) run on a lane of
// their own; or lane 0 runs all events in order. The same code runs on the CPU with one lane, port
// after port (egress ports of a safe slot are independent, so the order of their runs is free).
constexpr int KS = 32;  // ports of a switch whose slot runs from the LDS workspace
constexpr int KW = 64;  // ports of a switch whose slot runs from a device-memory workspace (a lane per port of wave64)
constexpr int LK = 512; // arrivals of a slot a safe slot takes
template <int K>
struct SwWorkT
{
    static constexpr int ports = K;
    Port pt[K];
    PortRun w[K];
    PortCfg pc[K]; // (a copy: no load of a port's configuration per event)
    Chain tc[K];   // the ports' pending TransmitCompletes' chains (copies, written back at the slot's end)
    int cnt[K], first[K];
    uint32_t abytes[K], ing0[K];
    int16_t mc[K][K]; // [egress][ingress]: the ingress's next arrival for the egress (-1: none)
    int16_t onext[LK];  // an arrival's next one of the same ingress and egress (-1: none)
    int64_t ats[LK];
    uint32_t acode[LK]; // its event code (the in-link's tx time and delay, its ordinal)
    int8_t own[LK];     // the egress port it belongs to in a safe slot (-1: none)
    uint8_t q[LK];
    int total, safe;
    unsigned act;
    unsigned nout;
};
typedef SwWorkT<KS> SwWork;     // in LDS, a wave's
typedef SwWorkT<KW> SwWorkWide; // in device memory, a wave's (switches of KS + 1 to KW ports)

#if defined(__HIP_DEVICE_COMPILE__)
#define RD_WAVE_SYNC()                                                                                                            \
    do                                                                                                                            \
    {                                                                                                                             \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");                                                                    \
        __builtin_amdgcn_wave_barrier();                                                                                          \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");                                                                    \
    } while (0)
#else
#define RD_WAVE_SYNC() \
    do                 \
    {                  \
    } while (0)
#endif

// arrival j's frame (of ingress p) in its link's ring (its code is in the workspace: k.acode[j])
template <typename WK>
RD_HD const LFrame RD_G *sw_arr(const Sim &s, const WK &k, int p, int j) { return s.msg + k.w[p].abase + (uint64_t)(j - k.first[p]); }

// egress port o's events of a safe slot (src/rdma/rstep.cuh:2806-3005)
#if defined(RD_PROF) && defined(__HIP_DEVICE_COMPILE__)
// (profiling builds: a wave's time in a part of the egress loop, counted once per wave)
#define RD_PROF_T(v) const long long v = wall_clock64()
#define RD_PROF_ADD(i, d)                                                                                                         \
    do                                                                                                                            \
    {                                                                                                                             \
        if ((int)(__builtin_ffsll((long long)__ballot(1)) - 1) == (int)__lane_id())                                               \
            at_add64(&s.ev[i], (unsigned long long)(d));                                                                          \
    } while (0)
#else
#define RD_PROF_T(v)
#define RD_PROF_ADD(i, d)
#endif
template <typename WK>
RD_HD void sw_egress_events(const Sim &s, SwSlot &c, WK &k, int o, int deg, int64_t t1)
{
    // the ingress ports with arrivals left for o (bit p: port p)
    uint64_t live = 0;
    for (int p = 0; p < deg; p++)
        if (k.mc[o][p] >= 0)
            live |= 1ull << p;
    for (;;)
    {
        RD_PROF_T(q0);
        // the next arrival for o: the ingress heads in port order, the earlier (time, chains) kept
        int bp = -1, kb = -1;
        for (uint64_t m = live; m; m &= m - 1)
        {
            const int p = __builtin_ctzll(m);
            const int j = k.mc[o][p];
            if (bp < 0)
            {
                bp = p, kb = j;
                continue;
            }
            bool b4;
            if (k.ats[j] != k.ats[kb])
                b4 = k.ats[j] < k.ats[kb];
            else
            {
                const LFrame RD_G *fa = sw_arr(s, k, p, j), *fb = sw_arr(s, k, bp, kb);
                b4 = before_q(s, k.ats[j], k.acode[j], fa->c, k.acode[kb], fb->c);
            }
            if (b4)
                bp = p, kb = j;
        }
        const int64_t ta = kb >= 0 ? k.ats[kb] : INT64_MAX, tc = k.pt[o].tc_at;
        const bool htc = tc >= 0 && tc < t1;
        RD_PROF_T(q1);
        RD_PROF_ADD(29, q1 - q0);
        if (ta == INT64_MAX && !htc)
            break;
        bool arr;
        const LFrame RD_G *fa = kb >= 0 ? sw_arr(s, k, bp, kb) : nullptr;
        const uint32_t acode = fa ? k.acode[kb] : 0;
        if (!htc || ta == INT64_MAX)
            arr = !htc;
        else if (ta != tc)
            arr = ta < tc;
        else
            arr = before_q(s, ta, acode, fa->c, 0, c.tc[o]);
        c.ord = 0;
        c.ep = o;
        c.cur_ok = false;
        if (!arr)
        {
            c.now = tc;
            c.cur_src = &c.tc[o];
            c.cur_code = 0;
            sw_event_chain(c);
            k.pt[o].tc_at = -1;
            port_pump(c, o);
            RD_PROF_T(q2);
            RD_PROF_ADD(30, q2 - q1);
            continue;
        }
        k.mc[o][bp] = k.onext[kb];
        if (k.onext[kb] < 0)
            live &= ~(1ull << bp);
        c.now = ta;
        c.cur_src = &fa->c;
        c.cur_code = acode;
        const Pkt pk = pkt_get(&fa->p);
        if (pk.kind == K_PAUSE || pk.kind == K_RESUME)
        {
            if (pk.kind == K_PAUSE)
                k.pt[o].paused |= 1u << pk.q;
            else
            {
                k.pt[o].paused &= ~(1u << pk.q);
                port_pump(c, o);
            }
        }
        else
            sw_ingress(c, bp, pk, o, k.q[kb]);
        RD_PROF_T(q3);
        RD_PROF_ADD(31, q3 - q1);
    }
}

// One switch's slot by lane `lane` of `nlanes` (a wave on the GPU, one lane on the CPU); ports
// p = lane, lane + nlanes, ... are this lane's. Returns (on every lane) whether the switch is busy.
template <typename WK>
RD_HD bool switch_slot(const Sim &s, int x, int64_t slot, uint32_t *pend, int pend_n, WK &k, int lane, int nlanes)
{
    const SwInfo &si = s.si[x];
    const int deg = si.deg;
    const int64_t t1 = (slot + 1) * s.D;
    SwSlot c;
    sw_ctx(c, s, x, slot, pend, pend_n, k.pt, k.w);
    c.pc = k.pc;
    c.tc = k.tc;
#if defined(RD_PROF) && defined(__HIP_DEVICE_COMPILE__)
    long long pt0 = wall_clock64();
#endif
    // ports in, arrivals counted
    for (int p = lane; p < deg; p += nlanes)
    {
        k.pc[p] = s.pc[si.pbase + p];
        k.pt[p] = s.ps[si.pbase + p];
        k.tc[p] = s.p_tc[si.pbase + p];
        k.ing0[p] = k.pt[p].ing;
        k.abytes[p] = 0;
        sw_port_open(s, c, p, slot);
        k.cnt[p] = k.w[p].an;
    }
    if (lane == 0)
    {
        k.act = 0;
        k.nout = 0;
    }
    RD_WAVE_SYNC();
#if defined(RD_PROF) && defined(__HIP_DEVICE_COMPILE__)
    long long pt1 = wall_clock64();
#endif
    int total = 0;
    for (int p = 0; p < deg; p++)
        total += k.cnt[p];
    // routes and times of the arrivals (a safe slot's candidates), per ingress port
    if (total <= LK)
        for (int p = lane; p < deg; p += nlanes)
        {
            int f0 = 0;
            for (int r = 0; r < p; r++)
                f0 += k.cnt[r];
            k.first[p] = f0;
            for (int i = 0; i < k.cnt[p]; i++)
            {
                const int j = f0 + i;
                const LFrame RD_G *f = s.msg + k.w[p].abase + i;
                k.ats[j] = coh_load(&f->ts);
                const Pkt pk = pkt_get(&f->p);
                k.acode[j] = ccode(port_in_tx(s, k.w[p], pk.size) + k.w[p].idelay, pk.ord);
                int q = 0;
                if (pk.kind == K_PAUSE || pk.kind == K_RESUME)
                    k.own[j] = (int8_t)p;
                else
                {
                    const int out = sw_route(s, *c.P, x, pk, &q);
                    k.own[j] = (int8_t)out;
                    if (out >= 0 && q != 0)
                        k.abytes[p] += pk.size;
                }
                k.q[j] = (uint8_t)q;
            }
            for (int o = 0; o < deg; o++)
                k.mc[o][p] = -1;
            for (int j = f0 + k.cnt[p] - 1; j >= f0; j--)
            {
                const int o = k.own[j];
                if (o < 0)
                    continue;
                k.onext[j] = k.mc[o][p];
                k.mc[o][p] = (int16_t)j;
            }
        }
    RD_WAVE_SYNC();
    // is the slot safe? (every lane decides alike)
    bool safe = !s.serial && total <= LK;
    if (safe)
    {
        uint64_t sum = 0;
        for (int p = 0; p < deg && safe; p++)
        {
            sum += k.abytes[p];
            safe = !k.pt[p].hdrm && !k.pt[p].mmu_paused;
        }
        // (GetPfcThreshold is uint32_t: with less buffer than headroom and reserves the difference
        // wraps to about 2^32 and stays there, so such a switch never pauses)
        const int64_t amax = (int64_t)c.sw.shared_used + (int64_t)sum;
        const int64_t room = (int64_t)(uint32_t)(c.P->buffer_size - c.sw.total_hdrm - c.sw.total_rsrv) - amax;
        const int64_t th_min = room > 0 ? room >> c.sw.max_shift : -1;
        for (int p = 0; p < deg && safe; p++)
        {
            const int64_t su = (int64_t)k.pt[p].ing + k.abytes[p] - RESERVE;
            safe = (su > 0 ? su : 0) < th_min;
        }
    }
#ifdef RD_PROF
    if (lane == 0)
        at_add64(&s.ev[safe ? 36 : 37], 1ull); // (profiling builds: safe and serial switch slots)
#endif
#if defined(RD_PROF) && defined(__HIP_DEVICE_COMPILE__)
    long long pt2 = wall_clock64();
#endif
    if (safe)
    {
        c.safe = true;
        for (int o = lane; o < deg; o += nlanes)
            sw_egress_events(s, c, k, o, deg, t1);
    }
    else if (lane == 0)
    {
        sw_serial_events(s, c, deg, t1);
        s.sw[x] = c.sw;
    }
    RD_WAVE_SYNC();
#if defined(RD_PROF) && defined(__HIP_DEVICE_COMPILE__)
    long long pt3 = wall_clock64();
#endif
    // ports out
    unsigned long long nout = 0;
    bool act = false;
    for (int p = lane; p < deg; p += nlanes)
    {
        act |= sw_port_close(s, c, p, &nout);
        s.ps[si.pbase + p] = k.pt[p];
        s.p_tc[si.pbase + p] = k.tc[p];
    }
    if (act)
        at_add32(&k.act, 1u);
    if (nout)
        at_add32(&k.nout, (unsigned)nout);
    RD_WAVE_SYNC();
    if (lane == 0)
    {
        if (safe)
        {
            // shared_used = sum over ports of max(0, ing - reserve), which changed with ing only
            // (modulo 2^32, as the counter)
            uint32_t d = 0;
            for (int p = 0; p < deg; p++)
            {
                const uint32_t a = k.ing0[p] > RESERVE ? k.ing0[p] - RESERVE : 0, b = k.pt[p].ing > RESERVE ? k.pt[p].ing - RESERVE : 0;
                d += b - a;
            }
            Switch sw = s.sw[x];
            sw.shared_used += d;
            s.sw[x] = sw;
        }
        if (k.nout)
            at_add64(&s.ev[EV_PKT], (unsigned long long)k.nout);
    }
    const bool busy = k.act != 0;
    RD_WAVE_SYNC(); // (the workspace is reused by the next switch)
#if defined(RD_PROF) && defined(__HIP_DEVICE_COMPILE__)
    // (profiling builds: the phases' time, summed over switch slots: ports in, routes, events, ports out;
    // and the arrivals)
    if (lane == 0)
    {
        const long long pt4 = wall_clock64();
        at_add64(&s.ev[24], (unsigned long long)(pt1 - pt0));
        at_add64(&s.ev[25], (unsigned long long)(pt2 - pt1));
        at_add64(&s.ev[26], (unsigned long long)(pt3 - pt2));
        at_add64(&s.ev[27], (unsigned long long)(pt4 - pt3));
        at_add64(&s.ev[28], (unsigned long long)total);
    }
#endif
    return busy;
}

// does the switch have events in this slot? (busy last slot, or arrivals; src/rdma/rstep.cuh:3859-3881)
RD_HD bool switch_has_input(const Sim &s, int x, int64_t slot)
{
    const SwInfo &si = s.si[x];
    for (int p = 0; p < si.deg; p++)
    {
        uint64_t base;
        if (bucket(s, s.pc[si.pbase + p].link_in, slot, &base) > 0)
            return true;
    }
    return false;
}

// The end of a world's slot, after all its entities ran it (src/rdma/rstep.cuh:3451-3585 without
// windows): finish when every flow is done or at the stop; with nothing sent or queued anywhere,
// skip to the slot of the earliest host event or pending arrival; with none, the run is quiescent.
RD_HD void slot_end(const Sim &s, const World &wd, WorldRun &r, uint32_t *pend, unsigned act, unsigned long long wake)
{
    // (the ticket and the pending-slot flags are read and written coherently: every workgroup of
    // the team reads the ticket after the barrier, and every entity sets flags)
    const int64_t slot = coh_load(&r.slot);
    // (an atomic read: the lanes' completions were counted by atomics, in L2)
    if (at_add32(&r.flows_done, 0u) >= (uint32_t)wd.nf || slot + 1 >= wd.max_slots)
        coh_store(&r.finished, 1);
    coh_store(pend + slot % wd.pend_n, 0u);
    int64_t next = slot + 1;
    if (act == 0)
    {
        int64_t ws = wake != ~0ull ? (int64_t)(wake / (unsigned long long)s.D) : INT64_MAX;
        for (int i = 1; i < wd.pend_n; i++)
            if (coh_load(pend + (slot + i) % wd.pend_n))
            {
                ws = slot + i < ws ? slot + i : ws;
                break;
            }
        if (ws != INT64_MAX && ws > next)
            next = ws < wd.max_slots ? ws : wd.max_slots;
        if (ws == INT64_MAX)
            coh_store(&r.finished, 1);
    }
    coh_store(&r.slots_run, coh_load(&r.slots_run) + 1);
    coh_store(&r.slot, next);
}

} // namespace rd
