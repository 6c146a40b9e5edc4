// TEST INFRASTRUCTURE ONLY -- not part of the product.
//
// A sequential CPU restatement of TwiNS's TCP packet model (the `twins` simulator): the same
// slot-synchronous execution, event order, integer picosecond time and floating-point
// expressions, written as plain single-threaded C++ (hosts one after another, then switches, then
// the slot ticket). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline run it, as
// the checker and as the CPU timing baseline; the product (twins_amd/) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the reference
// repository). What pins it: the stored sequential-ns-3 outputs (data/ns3/ref) that the reference
// reproduces, which the tests compare against column for column; the ECMP hash columns of the
// same files; libm's pow (glibc's FMA variant on this host, the routine ns-3 calls); x87 long
// double for Time::GetSeconds; L'Ecuyer's published MRG32k3a.
//
// The model is the reference's "general" form throughout (sent list of items, receiver blocks,
// an RTT history list); the reference's lean form is an equivalent compaction of it
// (validation/check_all.sh:11 checks both against ns-3).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace orc
{

// ---------------------------------------------------------------- constants (src/model.h:22-49)
constexpr int64_t SLOT_PS = 1000000, PROP_PS = 1000000, PS_PER_BYTE = 80;
constexpr uint32_t MSS = 1460;
constexpr int CTRL_FRAME = 42, SYN_FRAME = 46;
constexpr int MAX_TX = 300, QDISC_CAP = 200, RED_K = 30, MAX_ACTIVE = 256, MAX_DUE = 16;
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
enum { TOPO_FATTREE = 0, TOPO_LEAFSPINE = 1, TOPO_GRAPH = 2 };

constexpr int MAX_RUNS = 25, MAX_CHAIN = 2 * MAX_RUNS + 2;
static int RUNS = 9; // recorded runs of an event-order chain (9: `twins`, 25: `twins128`)

// counters, numbered as the reference prints them (src/main.cu:1066-1069)
static unsigned long long ev[32];
// packet-events: frames started on a link (each is one link traversal)
static unsigned long long tx_frames;

// ---------------------------------------------------------------- ns-3 Hash32 (src/ns3hash.h:8-47)
// MurmurHash3_x86_32 (Austin Appleby's published algorithm), seed 0x8BADF00D as ns-3 uses it.
static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static uint32_t hash32(const uint8_t *p, int n)
{
    uint32_t h = 0x8BADF00Du;
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    int i = 0;
    for (; i + 4 <= n; i += 4)
    {
        uint32_t k;
        memcpy(&k, p + i, 4); // little-endian host
        h ^= rotl32(k * c1, 15) * c2;
        h = rotl32(h, 13) * 5 + 0xe6546b64u;
    }
    uint32_t k = 0;
    for (int j = n - i - 1; j >= 0; j--)
        k = (k << 8) | p[i + j];
    if (n - i)
        h ^= rotl32(k * c1, 15) * c2;
    h ^= (uint32_t)n;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// ---------------------------------------------------------------- MRG32k3a (src/rng.h:29-159)
// L'Ecuyer, Simard, Chen, Kelton (2002), RngStreams, in doubles as ns-3's RngStream does.
struct Mrg
{
    double s[6];
};
static double mrg_u01(Mrg &g)
{
    const double m1 = 4294967087.0, m2 = 4294944443.0;
    double p1 = 1403580.0 * g.s[1] - 810728.0 * g.s[0];
    int32_t k = (int32_t)(p1 / m1);
    p1 -= k * m1;
    if (p1 < 0.0)
        p1 += m1;
    g.s[0] = g.s[1], g.s[1] = g.s[2], g.s[2] = p1;
    double p2 = 527612.0 * g.s[5] - 1370589.0 * g.s[3];
    k = (int32_t)(p2 / m2);
    p2 -= k * m2;
    if (p2 < 0.0)
        p2 += m2;
    g.s[3] = g.s[4], g.s[4] = g.s[5], g.s[5] = p2;
    const double norm = 1.0 / 4294967088.0;
    return p1 > p2 ? (p1 - p2) * norm : (p1 - p2 + m1) * norm;
}
// UniformRandomVariable::GetInteger(0, n - 1)
static uint32_t mrg_integer(Mrg &g, uint32_t n)
{
    double u = mrg_u01(g);
    double lo = 0.0, hi = (double)(n - 1) + 1.0;
    return (uint32_t)(lo + u * (hi - lo));
}
// RngStream(seed, stream 2^63 + s, substream run): the seed vector advanced by
// A^(stream * 2^127 + run * 2^76), per component, with exact modular matrix powers.
typedef unsigned __int128 u128;
struct M3
{
    uint64_t a[3][3];
};
static M3 m3_mul(const M3 &x, const M3 &y, uint64_t m)
{
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            u128 v = 0;
            for (int k = 0; k < 3; k++)
                v += (u128)x.a[i][k] * y.a[k][j] % m;
            r.a[i][j] = (uint64_t)(v % m);
        }
    return r;
}
static void m3_apply(const M3 &x, uint64_t *v, uint64_t m)
{
    uint64_t r[3];
    for (int i = 0; i < 3; i++)
    {
        u128 s = 0;
        for (int k = 0; k < 3; k++)
            s += (u128)x.a[i][k] * v[k] % m;
        r[i] = (uint64_t)(s % m);
    }
    memcpy(v, r, sizeof r);
}
// v = A^(n * 2^e) v
static void m3_jump(M3 a, uint64_t m, int e, uint64_t n, uint64_t *v)
{
    for (int i = 0; i < e; i++)
        a = m3_mul(a, a, m);
    while (n)
    {
        if (n & 1)
            m3_apply(a, v, m);
        a = m3_mul(a, a, m);
        n >>= 1;
    }
}
static Mrg mrg_stream(uint32_t seed, uint64_t stream, uint64_t run)
{
    const uint64_t m1 = 4294967087ull, m2 = 4294944443ull;
    M3 a1 = {{{0, 1, 0}, {0, 0, 1}, {m1 - 810728, 1403580, 0}}};
    M3 a2 = {{{0, 1, 0}, {0, 0, 1}, {m2 - 1370589, 0, 527612}}};
    uint64_t v1[3] = {seed, seed, seed}, v2[3] = {seed, seed, seed};
    uint64_t target = (1ull << 63) + stream;
    m3_jump(a1, m1, 127, target, v1);
    m3_jump(a2, m2, 127, target, v2);
    m3_jump(a1, m1, 76, run, v1);
    m3_jump(a2, m2, 76, run, v2);
    Mrg g;
    for (int i = 0; i < 3; i++)
        g.s[i] = (double)v1[i], g.s[3 + i] = (double)v2[i];
    return g;
}

// ---------------------------------------------------------------- ns-3 time and pow
// Time::GetSeconds() at picosecond resolution (src/nsmath.h:242-300): the Q64.64 product with
// int64x64_t::Invert(10^12), then int64x64_t::GetDouble(), which sums the high and the low word in
// x87 extended precision. Invert(10^12) = floor(2^128 / 10^12) + 1: ns-3 adds one because
// 10^12 * floor(2^128 / 10^12) < 2^128. This host's long double is the same x87 format.
static double ns_seconds(int64_t ps)
{
    const u128 inv = ~(u128)0 / 1000000000000ull + 1; // (2^128 - 1) / 10^12 = floor(2^128 / 10^12)
    bool neg = ps < 0;
    u128 a = neg ? (u128)(-(__int128)ps) : (u128)ps;
    u128 v = a * (uint64_t)(inv >> 64) + ((a * (uint64_t)inv) >> 64);
    long double hi = (long double)(uint64_t)(v >> 64);
    long double lo = (long double)(uint64_t)v / 18446744073709551616.0L;
    long double r = hi;
    r += lo;
    return (double)(neg ? -r : r);
}
// pow: libm's, the routine ns-3's TcpCubic calls (glibc selects its FMA variant on this host)
static double ns_pow(double x, double y) { return std::pow(x, y); }

// ---------------------------------------------------------------- event-order chains
// (src/step.cuh:58-159) A chain holds the leads from an event back towards its root as
// run-length (code, count) pairs, then the root key in two words. Codes grow with the lead.
struct Chain
{
    uint16_t w[MAX_CHAIN];
};
static inline uint16_t code_tc(int len) { return (uint16_t)len; }
static inline uint16_t code_arrival(int len) { return (uint16_t)(2048 + len); }
static inline uint16_t code_lead(int64_t lead_ps)
{
    int64_t c = 4096 + lead_ps / 16000;
    return (uint16_t)(c < 0xFFFE ? c : 0xFFFE);
}
constexpr uint16_t CODE_NOW = 1, CODE_STEP = 2;

static void chain_root(Chain &c, int host, int conn)
{
    memset(&c, 0, sizeof c);
    c.w[2 * RUNS] = (uint16_t)(0x8000 | (host + 1));
    c.w[2 * RUNS + 1] = (uint16_t)conn;
}
static void chain_root_msg(Chain &c, int msg)
{
    memset(&c, 0, sizeof c);
    c.w[2 * RUNS] = (uint16_t)(1 + (msg >> 16));
    c.w[2 * RUNS + 1] = (uint16_t)msg;
}
// child = [code] + parent (run-length: a repeated code extends the first run)
static Chain chain_child(uint16_t code, const Chain &p)
{
    Chain c;
    memset(&c, 0, sizeof c);
    if (p.w[0] == code)
    {
        c = p;
        if (c.w[1] < 0xFFFF)
            c.w[1]++;
        return c;
    }
    c.w[0] = code;
    c.w[1] = 1;
    for (int i = 2; i < 2 * RUNS; i++)
        c.w[i] = p.w[i - 2];
    c.w[2 * RUNS] = p.w[2 * RUNS];
    c.w[2 * RUNS + 1] = p.w[2 * RUNS + 1];
    return c;
}
// < 0: a runs first; > 0: b does; 0: undecided
static int chain_cmp(const Chain &A, const Chain &B)
{
    const uint16_t *a = A.w, *b = B.w;
    int ia = 0, ib = 0;
    uint32_t ra = a[1], rb = b[1];
    for (;;)
    {
        uint16_t ca = a[2 * ia], cb = b[2 * ib];
        if (ca == 0 && cb != 0 && a[2 * RUNS] != 0)
            return -1;
        if (cb == 0 && ca != 0 && b[2 * RUNS] != 0)
            return 1;
        if (ca == 0 || cb == 0)
            break;
        if (ca != cb)
            return ca > cb ? -1 : 1;
        uint32_t m = std::min(ra, rb);
        ra -= m, rb -= m;
        if (ra == 0)
        {
            if (++ia == RUNS)
                break;
            ra = a[2 * ia + 1];
        }
        if (rb == 0)
        {
            if (++ib == RUNS)
                break;
            rb = b[2 * ib + 1];
        }
    }
    uint32_t xa = (uint32_t)a[2 * RUNS] << 16 | a[2 * RUNS + 1], xb = (uint32_t)b[2 * RUNS] << 16 | b[2 * RUNS + 1];
    if (xa == 0 || xb == 0 || xa == xb)
        return 0;
    return xa < xb ? -1 : 1;
}

// ---------------------------------------------------------------- frames (src/model.h:76-88)
struct Frame
{
    int64_t ts;
    uint32_t flow, sa;
    uint16_t len, dst;
    uint8_t flags, to_sender;
    uint16_t up;
    Chain ch;
};
static inline uint32_t payload(const Frame &p) { return (p.flags & F_DATA) ? p.len - 42u : 0u; }

// the link messages: per ring slot (t mod 4) and link, the frames that started in slot `tag`
struct Msg
{
    int64_t tag = -1;
    std::vector<Frame> fr;
};

// ---------------------------------------------------------------- the queue trace (--qtrace)
// every enqueue at qt_from <= now < qt_until: time, queue id, the frame, the queue-disc length after
// it (src/step.cuh:266-279); written by write_qtrace as src/main.cu:1077-1100 does
struct QRec
{
    int64_t ts;
    int queue;
    uint32_t flow, sa;
    int len, qlen;
    uint8_t flags, to_sender;
};
static bool qt_on = false;
static int64_t qt_from = 0, qt_until = 0;
static std::vector<QRec> qtrace;

// ---------------------------------------------------------------- egress FIFO (src/step.cuh:167-317)
struct Queue
{
    std::vector<Frame> ring; // capacity: the largest queue limit + 2
    int id = 0;              // the reference's queue id: host h, or H + switch * k + port
    int head = 0, size = 0;
    int64_t link_free = 0, last_start = 0;
    Chain last_chain{};
    int red_k = RED_K, qcap = QDISC_CAP;
};

struct Fifo
{
    Queue *q;
    int64_t woke = -1, idle_start = -1, slot_end = 0;
    std::vector<Frame> *out = nullptr;

    // start every queued frame whose transmission begins at or before `until`, within the slot
    void advance(int64_t until)
    {
        Queue &Q = *q;
        const int cap = (int)Q.ring.size();
        while (Q.size > 0 && (int)out->size() < MAX_TX)
        {
            Frame &f = Q.ring[Q.head];
            int64_t start = std::max(f.ts, Q.link_free);
            if (start > until || start >= slot_end)
                return;
            int64_t end = start + (int64_t)f.len * PS_PER_BYTE;
            Frame o = f;
            // the start event: the previous frame's TransmitComplete if the frame waited,
            // else its own enqueue
            if (f.ts <= Q.link_free && Q.link_free > Q.last_start)
                o.ch = chain_child(code_tc((int)((Q.link_free - Q.last_start) / PS_PER_BYTE)), Q.last_chain);
            else
                idle_start = start;
            Q.last_chain = o.ch;
            Q.last_start = start;
            o.ts = end + PROP_PS;
            out->push_back(o);
            tx_frames++;
            Q.link_free = end;
            Q.head = (Q.head + 1) % cap;
            Q.size--;
        }
        if (Q.size > 0 && (int)out->size() >= MAX_TX)
            ev[4]++;
    }

    // RED (MinTh = MaxTh = K, QW = 1) on the queue-disc length, which leaves out the frame in
    // ns-3's one-packet device queue unless the Wake that refills it has not run yet
    bool enqueue(const Frame &f, int64_t now)
    {
        advance(now);
        Queue &Q = *q;
        bool refilled = Q.last_start < now || (woke == now && idle_start == now);
        int nq = (Q.size > 0 && Q.link_free > now && refilled) ? Q.size - 1 : Q.size;
        Frame g = f;
        g.ts = now;
        if (nq >= Q.red_k && nq > 1)
        {
            if (g.flags & F_ECT)
            {
                g.flags |= F_CE;
                ev[1]++;
            }
            else
            {
                ev[0]++;
                return false;
            }
        }
        if (nq >= Q.qcap)
        {
            ev[0]++;
            return false;
        }
        if (qt_on && now < qt_until && now >= qt_from)
            qtrace.push_back({now, Q.id, g.flow, g.sa, g.len, nq + 1, g.flags, g.to_sender});
        const int cap = (int)Q.ring.size();
        if (Q.size >= cap)
        {
            ev[19]++;
            return false;
        }
        Q.ring[(Q.head + Q.size) % cap] = g;
        Q.size++;
        return true;
    }
};

// ---------------------------------------------------------------- connection state
// One record per connection (a connection carries one or more messages), both sides' fields
// (src/model.h:118-226). Sequence numbers are relative to the first data byte.
struct Conn
{
    int src = 0, dst = 0;
    uint32_t hash_data = 0, hash_ack = 0;
    // the sending side and its application
    int m_first = 0, m_cnt = 0, m_due = 0, m_wr = 0, m_dn = 0;
    uint8_t fin_on = 0, st = S_IDLE, cong = C_OPEN, fin = 0;
    uint32_t tail = 0, una = 0, nxt = 0, high = 0, cwnd = INIT_CWND, ssthresh = 0xFFFFFFFFu, dupacks = 0;
    // the sent list (TcpTxBuffer items: start, flags) and the RTT history, rings of it_cap entries
    uint32_t it_cap = 0, it_head = 0, it_n = 0;
    std::vector<uint32_t> it_start;
    std::vector<uint8_t> it_fl;
    uint32_t sack_b = 0, lost_b = 0, rtx_b = 0, sacked_n = 0, cand_n = 0, cand = 0, hisack = 0;
    uint8_t sack_seen = 0;
    uint32_t h_head = 0, h_n = 0;
    std::vector<uint32_t> h_seq, h_cnt;
    std::vector<int64_t> h_t;
    uint32_t recover = 0;
    uint8_t recover_active = 0, first_partial = 1, cwnd_limited = 0;
    uint32_t cwnd_cnt = 0, ack_rem = 0;
    // TcpCubic
    uint32_t cub_cnt = 0, cub_lastmax = 0, cub_origin = 0, cub_endseq = 0, cub_samples = 0, cub_ackcnt = 0, cub_tcpcwnd = 0;
    double cub_k = 0.0;
    int64_t cub_delaymin = INT64_MIN, cub_epoch = INT64_MIN, cub_roundstart = INT64_MIN, cub_lastack = INT64_MIN,
            cub_currrtt = INT64_MIN;
    uint8_t cub_found = 0;
    // ECN / DCTCP at the sender
    int64_t ecn_echo_seq = -1, ecn_cwr_seq = -1;
    uint8_t s_ce = 0, ecn_s = E_IDLE, next_seq_flag = 0;
    double alpha = 1.0;
    uint32_t acked_total = 0, acked_ecn = 0, next_seq = 0;
    // RTT estimator and timers of the sender
    int64_t srtt = -1, rttvar = 0, rto = MIN_RTO_PS, rto_at = -1, timing_time = 0, syn_at = -1;
    uint8_t syn_tries = 0;
    // the receiving side: out-of-order blocks [oo_s, oo_e), a ring of oo_cap
    uint32_t oo_cap = 0, ooo_n = 0, ooo_h = 0, rcv_nxt = 0, prior_rcv_nxt = 0;
    std::vector<uint32_t> oo_s, oo_e;
    uint8_t r_up = 0, r_est = 0, r_st = R_OPEN, r_last_n = 0, srv_syn_n = 0, delack_count = 0, ce_state = 0, prior_flag = 0,
            reserved = 0, ecn_r = E_IDLE;
    int64_t r_fin = -1, r_synack_t = -1, r_srtt = -1, r_rttvar = 0, r_last_at = -1, r_frx_at = -1, srv_syn_at = -1,
            delack_at = -1, ecn_ce_seq = -1;
    // event-order chains of the timers: delayed ACK (and the receiver's others), retransmission
    // (or SYN), the deferred SendPendingData (and the receiver's FIN+ACK retransmission)
    Chain dchain{}, rchain{}, schain{};
};

struct Due
{
    int64_t at;
    int flow, kind; // 0 rto, 1 delack, 2 syn / deferred send, 3 SYN-ACK retx, 4 LastAck, 5 FIN+ACK retx
};

struct Host
{
    Queue nic;
    uint8_t cc = CC_DCTCP;
    std::vector<int> act_snd, act_rcv;
    int pend_next = 0, pend_end = 0;
    int64_t wake = 0;
};

struct Topo
{
    int kind = TOPO_FATTREE;
    int k = 4, hk = 2, H = 0, S = 0, ecmp_n = 1, leaves = 0, spines = 0, hpl = 0;
    // topology files: a host's switch and port; per epoch, equal-cost port lists and down ports
    std::vector<int> host_sw, host_port;
    int n_fail = 0;
    std::vector<int64_t> fail_t;
    std::vector<uint32_t> rt_off;
    std::vector<uint8_t> rt_n, rt_ports, port_dn;
    std::vector<int> host_in, sw_in; // the link each host port / switch port receives on (-1: none)
};

struct Sim;

struct HostCtx
{
    Sim *s;
    Host *host;
    Fifo *nic;
    int h;
    bool dctcp, cubic;
    Chain echain; // the chain of the event being processed
    std::vector<Due> *due;
    int nd, id; // timers due in this slot: [id, nd) pending
    int64_t slot_end;
};

struct Sim
{
    Topo T;
    std::vector<Conn> F;
    // messages (the application's flows), numbered so that a connection's are consecutive
    std::vector<int> mg_conn;
    std::vector<uint32_t> mg_size, mg_end;
    std::vector<int64_t> mg_start, mg_est, mg_done;
    std::vector<int> pend_msg; // per host, message ids in (start, writes before starts) order
    std::vector<Host> hosts;
    std::vector<Queue> ports; // [S * k]
    std::vector<Mrg> sw_rng;  // per-packet ECMP streams (empty: per-flow ECMP)
    std::vector<Msg> msg;     // [4][L]
    int L = 0, M = 0;
    unsigned flows_done = 0;
    int64_t max_slots = 2000000;

    Msg &message(int64_t slot, int link) { return msg[(size_t)(slot & 3) * L + link]; }
    const std::vector<Frame> *frames(int64_t slot, int link)
    {
        Msg &m = message(slot, link);
        return m.tag == slot ? &m.fr : nullptr;
    }

    // ------------------------------------------------------------ sent list (src/step.cuh:409-975)
    static uint32_t at(const Conn &c, uint32_t i)
    {
        uint32_t x = c.it_head + i;
        return x >= c.it_cap ? x - c.it_cap : x;
    }
    static uint32_t beg(const Conn &c, uint32_t i) { return c.it_start[at(c, i)]; }
    static uint32_t end(const Conn &c, uint32_t i) { return i + 1 < c.it_n ? c.it_start[at(c, i + 1)] : c.high; }
    static uint8_t iflags(const Conn &c, uint32_t i) { return c.it_fl[at(c, i)]; }
    static bool cand_of(uint8_t fl) { return (fl & (IT_LOST | IT_RETX | IT_SACKED)) == IT_LOST; }

    // set item i's flags, keeping the byte counters, the sacked count and the NextSeg cursor
    static void it_set(Conn &c, uint32_t i, uint8_t nf)
    {
        uint32_t k = at(c, i);
        uint8_t of = c.it_fl[k];
        if (of == nf)
            return;
        uint32_t len = end(c, i) - c.it_start[k];
        uint8_t d = of ^ nf;
        if (d & IT_SACKED)
        {
            c.sack_b += (nf & IT_SACKED) ? len : 0u - len;
            c.sacked_n += (nf & IT_SACKED) ? 1u : 0u - 1u;
        }
        if (d & IT_LOST)
            c.lost_b += (nf & IT_LOST) ? len : 0u - len;
        if (d & IT_RETX)
            c.rtx_b += (nf & IT_RETX) ? len : 0u - len;
        if (cand_of(of) != cand_of(nf))
        {
            if (cand_of(nf))
            {
                c.cand_n++;
                if (i < c.cand)
                    c.cand = i;
            }
            else
                c.cand_n--;
        }
        c.it_fl[k] = nf;
    }
    static bool it_insert(Conn &c, uint32_t i, uint32_t start, uint8_t fl)
    {
        if (c.it_n >= c.it_cap)
        {
            ev[14]++;
            return false;
        }
        for (uint32_t j = c.it_n; j > i; j--)
        {
            c.it_start[at(c, j)] = c.it_start[at(c, j - 1)];
            c.it_fl[at(c, j)] = c.it_fl[at(c, j - 1)];
        }
        c.it_start[at(c, i)] = start;
        c.it_fl[at(c, i)] = fl;
        c.it_n++;
        if (c.cand >= i && c.cand > 0)
            c.cand++;
        if (cand_of(fl))
        {
            c.cand_n++;
            if (i < c.cand)
                c.cand = i;
        }
        return true;
    }
    static void it_remove_start(Conn &c, uint32_t i)
    {
        for (uint32_t j = i; j + 1 < c.it_n; j++)
        {
            c.it_start[at(c, j)] = c.it_start[at(c, j + 1)];
            c.it_fl[at(c, j)] = c.it_fl[at(c, j + 1)];
        }
        c.it_n--;
        if (c.cand > i)
            c.cand--;
    }
    static uint32_t it_lower(const Conn &c, uint32_t seq)
    {
        uint32_t lo = 0, hi = c.it_n;
        while (lo < hi)
        {
            uint32_t m = (lo + hi) / 2;
            if (beg(c, m) < seq)
                lo = m + 1;
            else
                hi = m;
        }
        return lo;
    }
    static uint32_t tb_sent(const Conn &c) { return c.it_n ? c.high - c.una : 0u; }
    static uint32_t inflight(const Conn &c) { return tb_sent(c) - (c.sack_b + c.lost_b) + c.rtx_b; }
    static void tb_add_reno_sack(Conn &c)
    {
        uint32_t i = 1 + c.sacked_n;
        while (i < c.it_n && (iflags(c, i) & IT_SACKED))
            i++;
        if (i < c.it_n)
        {
            it_set(c, i, iflags(c, i) | IT_SACKED);
            c.sack_seen = 1;
            c.hisack = beg(c, i);
        }
    }
    static void tb_mark_head_lost(Conn &c)
    {
        if (c.it_n == 0)
            return;
        it_set(c, 0, (uint8_t)((iflags(c, 0) & ~(IT_SACKED | IT_RETX)) | IT_LOST));
    }
    static void tb_all_lost(Conn &c)
    {
        for (uint32_t i = 0; i < c.it_n; i++)
            c.it_fl[at(c, i)] = IT_LOST;
        c.sack_b = 0;
        c.sacked_n = 0;
        c.rtx_b = 0;
        c.lost_b = tb_sent(c);
        c.cand_n = c.it_n;
        c.cand = 0;
        c.sack_seen = 0;
        c.hisack = 0;
    }
    // DiscardUpTo(ack): returns the bytes of the items removed whole
    static uint32_t tb_discard(Conn &c, uint32_t ack)
    {
        uint32_t delivered = 0, removed = 0;
        while (c.it_n > 0)
        {
            uint32_t k = at(c, 0), b = c.it_start[k], e = end(c, 0);
            if (ack >= e)
            {
                if (c.it_fl[k])
                    it_set(c, 0, 0);
                c.it_head = at(c, 1);
                c.it_n--;
                removed++;
                delivered += e - b;
            }
            else
            {
                if (ack > b)
                {
                    uint8_t fl = c.it_fl[k];
                    uint32_t off = ack - b;
                    if (fl & IT_SACKED)
                        c.sack_b -= off;
                    if (fl & IT_LOST)
                        c.lost_b -= off;
                    if (fl & IT_RETX)
                        c.rtx_b -= off;
                    c.it_start[k] = ack;
                }
                break;
            }
        }
        if (removed)
            c.cand = c.cand > removed ? c.cand - removed : 0;
        c.una = ack;
        if (c.it_n > 0 && (iflags(c, 0) & IT_SACKED))
        {
            it_set(c, 0, (uint8_t)(iflags(c, 0) & ~IT_SACKED));
            tb_add_reno_sack(c);
            tb_mark_head_lost(c);
        }
        if (c.hisack <= ack)
        {
            c.sack_seen = 0;
            c.hisack = 0;
        }
        return delivered;
    }
    static bool tb_next_seg(Conn &c, uint32_t *seq)
    {
        if (c.cand_n > 0)
        {
            for (uint32_t i = c.cand; i < c.it_n; i++)
            {
                if (!cand_of(iflags(c, i)))
                    continue;
                c.cand = i;
                uint32_t b = beg(c, i);
                if (!c.sack_seen || b < c.hisack)
                {
                    *seq = b;
                    return true;
                }
                break;
            }
        }
        if (c.tail > c.high)
        {
            *seq = c.high;
            return true;
        }
        return false;
    }
    static bool tb_head_lost(const Conn &c)
    {
        if (!c.sack_seen || c.una >= c.hisack || c.it_n == 0)
            return false;
        return iflags(c, 0) & IT_LOST;
    }
    static bool tb_retx_ends_at(const Conn &c, uint32_t x)
    {
        if (c.it_n == 0 || x <= c.una || x > c.high)
            return false;
        uint32_t lo = it_lower(c, x);
        if (lo == 0)
            return false; // (x would lie before the head)
        uint32_t i = lo - 1;
        if (end(c, i) != x)
            return false;
        uint8_t fl = iflags(c, i);
        return !(fl & IT_SACKED) && (fl & IT_RETX);
    }
    // GetTransmittedSegment for item i: merge with an unsacked successor of the same lost state,
    // split to the size sent, mark retransmitted
    static uint32_t tb_retransmit(Conn &c, uint32_t i, uint32_t num)
    {
        uint32_t len = end(c, i) - beg(c, i), sz = num;
        if (i + 1 < c.it_n)
        {
            uint8_t a = iflags(c, i), b = iflags(c, i + 1);
            uint32_t nlen = end(c, i + 1) - beg(c, i + 1);
            if (!(b & IT_SACKED) && (a & IT_LOST) == (b & IT_LOST))
                sz = std::min(sz, len + nlen);
            else
                sz = std::min(sz, len);
        }
        else
            sz = std::min(sz, len);
        if (sz > len)
        {
            uint8_t a = iflags(c, i), b = iflags(c, i + 1);
            if ((a ^ b) & IT_RETX)
            {
                if (a & IT_RETX)
                    it_set(c, i, a &= (uint8_t)~IT_RETX);
                else
                    it_set(c, i + 1, b &= (uint8_t)~IT_RETX);
            }
            if (cand_of(b))
                c.cand_n--;
            it_remove_start(c, i + 1);
            len = end(c, i) - beg(c, i);
        }
        if (sz < len)
            it_insert(c, i + 1, beg(c, i) + sz, iflags(c, i));
        uint8_t fl = iflags(c, i);
        if (!(fl & IT_RETX))
            it_set(c, i, fl | IT_RETX);
        return sz;
    }
    // CopyFromSequence(num, seq): the segment's size; new data become a new item
    static uint32_t tb_copy(Conn &c, uint32_t num, uint32_t seq, bool *retx)
    {
        for (;;)
        {
            uint32_t avail = c.tail > seq ? c.tail - seq : 0;
            uint32_t sz = std::min(num, avail);
            if (sz == 0)
                return 0;
            uint32_t sent_end = c.high;
            if (sent_end >= seq + sz)
            {
                uint32_t i = it_lower(c, seq);
                if (i >= c.it_n || beg(c, i) != seq)
                {
                    ev[14]++;
                    return 0;
                }
                *retx = true;
                return tb_retransmit(c, i, sz);
            }
            if (sent_end <= seq)
            {
                if (!it_insert(c, c.it_n, seq, 0))
                    return 0;
                c.high = seq + sz;
                *retx = false;
                return sz;
            }
            num = sent_end - seq;
        }
    }

    // ------------------------------------------------------------ RTT (src/step.cuh:977-1126)
    static uint32_t hat(const Conn &c, uint32_t j)
    {
        uint32_t x = c.h_head + j;
        return x >= c.it_cap ? x - c.it_cap : x;
    }
    static void rtt_history(Conn &c, uint32_t seq, uint32_t sz, bool retx, int64_t now)
    {
        if (!retx)
        {
            if (c.h_n >= c.it_cap)
            {
                ev[14]++;
                return;
            }
            uint32_t k = hat(c, c.h_n);
            c.h_seq[k] = seq;
            c.h_cnt[k] = sz;
            c.h_t[k] = now;
            c.h_n++;
            return;
        }
        for (uint32_t j = 0; j < c.h_n; j++)
        {
            uint32_t k = hat(c, j), n = c.h_cnt[k] & 0x7FFFFFFFu;
            if (seq >= c.h_seq[k] && seq < c.h_seq[k] + n)
            {
                c.h_cnt[k] = 0x80000000u | (seq + sz - c.h_seq[k]);
                return;
            }
        }
    }
    static void rtt_sample(Conn &c, int64_t m)
    {
        int64_t est = c.srtt, var = c.rttvar;
        if (est < 0)
        {
            est = m;
            var = m / 2;
        }
        else
        {
            int64_t delta = m - est;
            est = ((est << 3) + delta) >> 3;
            if (delta < 0)
                delta = -delta;
            var = ((var << 2) + delta - var) >> 2;
        }
        c.srtt = est;
        c.rttvar = var;
        int64_t v = 4 * var;
        int64_t r = est + std::max(v, CLOCK_GRAN_PS);
        c.rto = std::max(r, MIN_RTO_PS);
    }
    static void rtt_ack(Conn &c, uint32_t ack, int64_t now)
    {
        if (c.h_n == 0)
            return;
        uint32_t k = hat(c, 0);
        if (!(c.h_cnt[k] & 0x80000000u) && ack >= c.h_seq[k] + c.h_cnt[k])
            rtt_sample(c, now - c.h_t[k]);
        while (c.h_n > 0)
        {
            k = hat(c, 0);
            if (c.h_seq[k] + (c.h_cnt[k] & 0x7FFFFFFFu) > ack)
                break;
            c.h_head = hat(c, 1);
            c.h_n--;
        }
    }

    // ------------------------------------------------------------ host TCP (src/step.cuh:319-2261)
    void due_insert(HostCtx &x, int64_t at_, int f, int kind)
    {
        if (at_ >= x.slot_end)
            return;
        std::vector<Due> &d = *x.due;
        if (x.nd >= MAX_DUE)
        {
            ev[7]++;
            return;
        }
        int j = x.nd;
        while (j > x.id && d[j - 1].at > at_)
        {
            d[j] = d[j - 1];
            j--;
        }
        d[j] = {at_, f, kind};
        x.nd++;
    }
    static void timer_chain(HostCtx &x, Chain &c, int64_t lead) { c = chain_child(code_lead(lead), x.echain); }

    void emit(HostCtx &x, int f, uint8_t flags, uint32_t seq, uint32_t ack, uint16_t len, bool to_sender, int64_t now)
    {
        const Conn &c = F[f];
        Frame fr;
        fr.ts = now;
        fr.flow = (uint32_t)f;
        fr.sa = (flags & F_DATA) ? seq : ack;
        fr.len = len;
        fr.dst = (uint16_t)(to_sender ? c.src : c.dst);
        fr.flags = flags | (x.dctcp ? F_ECT : 0);
        fr.to_sender = to_sender ? 1 : 0;
        fr.up = (uint16_t)((to_sender ? c.hash_ack : c.hash_data) % (uint32_t)T.ecmp_n);
        fr.ch = x.echain;
        x.nic->enqueue(fr, now);
    }
    void send_ack(HostCtx &x, int f, uint32_t ack, bool force_ece, bool no_ece, int64_t now)
    {
        Conn &c = F[f];
        bool ece = false;
        if (force_ece)
            ece = true;
        else if (!no_ece && (c.ecn_r == E_CE_RCVD || c.ecn_r == E_SENDING_ECE))
        {
            ece = true;
            c.ecn_r = E_SENDING_ECE;
        }
        c.delack_at = -1;
        c.delack_count = 0;
        emit(x, f, F_ACK | (ece ? F_ECE : 0), 0, ack, CTRL_FRAME, true, now);
    }

    // TcpCubic (src/step.cuh:1215-1360)
    static uint32_t cub_u32(double d) { return (uint32_t)(uint64_t)(int64_t)d; }
    static void cub_hystart_reset(Conn &c, int64_t now)
    {
        c.cub_roundstart = c.cub_lastack = now;
        c.cub_endseq = c.st == S_EST ? c.high + 1 : 0;
        c.cub_currrtt = INT64_MIN;
        c.cub_samples = 0;
    }
    static uint32_t cub_update(Conn &c, uint32_t segs, int64_t now)
    {
        const double C = 0.4, BETA = 0.7;
        uint32_t seg_cwnd = c.cwnd / MSS;
        c.cub_ackcnt += segs;
        if (c.cub_epoch == INT64_MIN)
        {
            c.cub_epoch = now;
            c.cub_ackcnt = segs;
            c.cub_tcpcwnd = seg_cwnd;
            if (c.cub_lastmax <= seg_cwnd)
            {
                c.cub_k = 0.0;
                c.cub_origin = seg_cwnd;
            }
            else
            {
                c.cub_k = ns_pow((double)(c.cub_lastmax - seg_cwnd) / C, 1 / 3.);
                c.cub_origin = c.cub_lastmax;
            }
        }
        int64_t t = (int64_t)((uint64_t)now + (uint64_t)c.cub_delaymin - (uint64_t)c.cub_epoch);
        double ts = ns_seconds(t), k = c.cub_k;
        double offs = ts < k ? k - ts : ts - k;
        uint32_t delta = cub_u32(C * ns_pow(offs, 3));
        uint32_t target = ts < k ? c.cub_origin - delta : c.cub_origin + delta;
        uint32_t cnt = target > seg_cwnd ? seg_cwnd / (target - seg_cwnd) : 100 * seg_cwnd;
        if (c.cub_lastmax == 0 && cnt > 20)
            cnt = 20;
        uint32_t scale = cub_u32(8 * (1024 + BETA * 1024) / 3 / (1024 - BETA * 1024));
        delta = (seg_cwnd * scale) >> 3;
        while (c.cub_ackcnt > delta)
        {
            c.cub_ackcnt -= delta;
            c.cub_tcpcwnd++;
        }
        if (c.cub_tcpcwnd > seg_cwnd)
        {
            uint32_t max_cnt = seg_cwnd / (c.cub_tcpcwnd - seg_cwnd);
            if (cnt > max_cnt)
                cnt = max_cnt;
        }
        return cnt > 2 ? cnt : 2;
    }
    static void cub_pkts_acked(Conn &c, int64_t now)
    {
        const int64_t ACK_DELTA = 2000000000, DELAY_MIN = 4000000000, DELAY_MAX = 1000000000000, CUBIC_DELTA = 10000000000;
        if (c.cub_epoch != INT64_MIN && now - c.cub_epoch < CUBIC_DELTA)
            return;
        int64_t rtt = c.srtt;
        if (c.cub_delaymin == INT64_MIN || c.cub_delaymin > rtt)
            c.cub_delaymin = rtt;
        uint32_t cwnd = c.cwnd;
        if (!(cwnd <= c.ssthresh && cwnd >= 16 * MSS) || c.cub_found)
            return;
        if (now - c.cub_lastack <= ACK_DELTA)
        {
            c.cub_lastack = now;
            if (now - c.cub_roundstart > c.cub_delaymin)
                c.cub_found = 1;
        }
        if (c.cub_samples < 8)
        {
            if (c.cub_currrtt == INT64_MIN || c.cub_currrtt > rtt)
                c.cub_currrtt = rtt;
            c.cub_samples++;
        }
        else
        {
            int64_t dm = c.cub_delaymin;
            int64_t th = dm > DELAY_MAX ? DELAY_MAX : (dm < DELAY_MIN ? DELAY_MIN : dm);
            if (c.cub_currrtt > dm + th)
                c.cub_found = 1;
        }
        if (c.cub_found)
            c.ssthresh = cwnd;
    }
    static void cub_increase(Conn &c, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (!c.cwnd_limited)
            return;
        if (c.cwnd < c.ssthresh)
        {
            if (ack + 1 > c.cub_endseq)
                cub_hystart_reset(c, now);
            c.cwnd += segs * MSS;
            segs = 0;
        }
        if (c.cwnd >= c.ssthresh && segs > 0)
        {
            c.cub_cnt += segs;
            uint32_t cnt = cub_update(c, segs, now);
            if (c.cub_cnt >= cnt)
            {
                c.cwnd += MSS;
                c.cub_cnt -= cnt;
            }
        }
    }
    static uint32_t cub_ssthresh(Conn &c)
    {
        const double BETA = 0.7;
        uint32_t seg_cwnd = c.cwnd / MSS;
        if (seg_cwnd < c.cub_lastmax)
            c.cub_lastmax = cub_u32((seg_cwnd * (1 + BETA)) / 2);
        else
            c.cub_lastmax = seg_cwnd;
        c.cub_epoch = INT64_MIN;
        uint32_t th = cub_u32(seg_cwnd * BETA);
        return (th > 2 ? th : 2) * MSS;
    }
    static void cub_loss(Conn &c, int64_t now)
    {
        c.cub_origin = 0;
        c.cub_k = 0;
        c.cub_ackcnt = 0;
        c.cub_tcpcwnd = 0;
        c.cub_delaymin = INT64_MIN;
        c.cub_found = 0;
        cub_hystart_reset(c, now);
    }

    // GetSsThresh of the host's TCP (src/step.cuh:1057-1066)
    static uint32_t ssthresh_for(HostCtx &x, Conn &c, uint32_t bytes_in_flight)
    {
        if (x.cubic)
            return cub_ssthresh(c);
        if (x.dctcp)
            return (uint32_t)((1.0 - c.alpha / 2.0) * c.cwnd);
        uint32_t half = bytes_in_flight / 2;
        return half > 2 * MSS ? half : 2 * MSS;
    }
    // PktsAcked (src/step.cuh:1362-1393)
    static void pkts_acked(HostCtx &x, Conn &c, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (x.cubic)
        {
            cub_pkts_acked(c, now);
            return;
        }
        if (!x.dctcp)
            return;
        c.acked_total += segs * MSS;
        if (c.ecn_s == E_ECE_RCVD)
            c.acked_ecn += segs * MSS;
        if (!c.next_seq_flag)
        {
            c.next_seq = c.nxt;
            c.next_seq_flag = 1;
        }
        if (ack >= c.next_seq)
        {
            double m = c.acked_total > 0 ? (double)c.acked_ecn / c.acked_total : 0.0;
            c.alpha = std::fma(1.0 - DCTCP_G, c.alpha, DCTCP_G * m);
            c.next_seq = c.nxt;
            c.acked_total = 0;
            c.acked_ecn = 0;
        }
    }
    // IncreaseWindow (src/step.cuh:1395-1455)
    static void increase_window(HostCtx &x, Conn &c, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (x.cubic)
        {
            cub_increase(c, segs, ack, now);
            return;
        }
        uint32_t cwnd = c.cwnd, ssth = c.ssthresh;
        if (x.dctcp)
        {
            if (cwnd < ssth)
            {
                if (segs >= 1)
                {
                    uint64_t n = (uint64_t)cwnd + (uint64_t)segs * MSS;
                    c.cwnd = (uint32_t)(n < ssth ? n : ssth);
                }
            }
            else
            {
                uint32_t w = cwnd / MSS;
                if (w == 0)
                    w = 1;
                uint32_t cnt = c.cwnd_cnt;
                if (cnt >= w)
                {
                    cnt = 0;
                    cwnd += MSS;
                }
                cnt += segs;
                if (cnt >= w)
                {
                    uint32_t delta = cnt / w;
                    cnt -= delta * w;
                    cwnd += delta * MSS;
                }
                c.cwnd = cwnd;
                c.cwnd_cnt = cnt;
            }
            return;
        }
        if (cwnd < ssth && segs >= 1)
        {
            cwnd += MSS;
            segs -= 1;
        }
        if (cwnd >= ssth && segs > 0)
        {
            double adder = (double)MSS * MSS / cwnd;
            if (adder < 1.0)
                adder = 1.0;
            cwnd += (uint32_t)adder;
        }
        c.cwnd = cwnd;
    }

    // SendDataPacket (src/step.cuh:1128-1171)
    uint32_t send_segment(HostCtx &x, int f, uint32_t seq, uint32_t max_size, int64_t now)
    {
        Conn &c = F[f];
        bool retx = false;
        uint32_t len = tb_copy(c, max_size, seq, &retx);
        if (len == 0)
            return 0;
        uint8_t flags = F_DATA;
        if (!retx && c.ecn_s == E_ECE_RCVD && c.ecn_echo_seq > c.ecn_cwr_seq)
        {
            c.ecn_s = E_CWR_SENT;
            c.ecn_cwr_seq = seq;
            flags |= F_CWR;
        }
        if ((c.fin & FIN_CLOSE) && seq + len == c.tail)
        {
            flags |= F_FIN;
            c.fin |= FIN_SENT;
        }
        emit(x, f, flags, seq, 0, (uint16_t)(len + 42), false, now);
        if (retx)
        {
            ev[2]++;
            ev[8 + c.cong]++;
        }
        if (c.rto_at < 0)
        {
            c.rto_at = now + c.rto;
            timer_chain(x, c.rchain, c.rto);
        }
        c.cwnd_limited = c.cwnd < inflight(c) + MSS;
        rtt_history(c, seq, len, retx, now);
        if (seq + len > c.high)
            c.high = seq + len;
        ev[13] = std::max<unsigned long long>(ev[13], c.it_n);
        return len;
    }
    // SendPendingData (src/step.cuh:1173-1200)
    void send_pending(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        if (c.st != S_EST)
            return;
        for (int guard = 0; guard < 4096; guard++)
        {
            if (c.cong == C_OPEN && (c.fin & (FIN_SENT | FIN_ACKED)) == FIN_SENT)
                return;
            uint32_t pipe = inflight(c), cwnd = c.cwnd;
            uint32_t avail = cwnd > pipe ? cwnd - pipe : 0;
            if (avail == 0)
                return;
            uint32_t seq;
            if (!tb_next_seg(c, &seq))
                return;
            uint32_t remaining = c.tail - seq;
            if (avail < MSS && remaining > avail)
                return;
            c.nxt = seq;
            uint32_t len = send_segment(x, f, seq, avail < MSS ? avail : MSS, now);
            if (len == 0)
                return;
            c.nxt = seq + len;
        }
    }
    void do_retransmit(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        uint32_t seq;
        if (!tb_next_seg(c, &seq))
            seq = c.una;
        if (c.una >= c.tail)
            return;
        c.nxt = seq;
        send_segment(x, f, seq, MSS, now);
    }
    // NewAck (src/step.cuh:1457-1479)
    void new_ack(HostCtx &x, int f, uint32_t ack, bool reset_rto, int64_t now)
    {
        Conn &c = F[f];
        if (reset_rto)
        {
            int64_t est = c.srtt >= 0 ? c.srtt : INIT_RTT_PS, v = 4 * c.rttvar;
            int64_t r = est + std::max(v, CLOCK_GRAN_PS);
            c.rto = std::max(r, MIN_RTO_PS);
            c.rto_at = now + c.rto;
            timer_chain(x, c.rchain, c.rto);
        }
        if (ack >= c.tail && !(c.fin & FIN_SENT))
            c.rto_at = -1;
        if (ack > c.nxt)
            c.nxt = ack;
    }
    void enter_recovery(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        tb_add_reno_sack(c);
        tb_mark_head_lost(c);
        c.recover = c.high;
        c.recover_active = 1;
        c.cong = C_RECOVERY;
        c.ssthresh = ssthresh_for(x, c, inflight(c) + MSS);
        c.cwnd = c.ssthresh;
        send_segment(x, f, c.una, MSS, now);
    }
    void messages_acked(int f, int64_t now)
    {
        Conn &c = F[f];
        while (c.m_dn < c.m_wr && c.una >= mg_end[c.m_first + c.m_dn])
        {
            mg_done[c.m_first + c.m_dn] = now;
            c.m_dn++;
            flows_done++;
        }
        if (c.m_dn == c.m_cnt && (!c.fin_on || (c.fin & FIN_ACKED)))
        {
            c.st = S_DONE;
            c.rto_at = -1;
        }
    }
    // ReceivedAck + ProcessAck (src/step.cuh:1515-1692)
    void on_ack(HostCtx &x, int f, uint32_t ack, bool ece, int64_t now)
    {
        Conn &c = F[f];
        if (c.st != S_EST)
            return;
        uint32_t old = c.una;
        if (ack < old)
            return;
        rtt_ack(c, ack, now);
        bool advancing = ack > old;
        uint32_t delivered = 0;
        if (advancing)
            delivered = tb_discard(c, ack);
        if (c.cong == C_CWR && ack > c.recover)
        {
            c.cong = C_OPEN;
            c.cwnd = c.ssthresh;
        }
        if (advancing && x.dctcp && ece)
        {
            if (c.ecn_echo_seq < (int64_t)ack)
            {
                c.ecn_echo_seq = ack;
                c.ecn_s = E_ECE_RCVD;
                if (c.cong != C_CWR)
                {
                    c.ssthresh = ssthresh_for(x, c, inflight(c));
                    c.cwnd = c.ssthresh;
                    c.cong = C_CWR;
                    c.recover = c.high;
                }
            }
        }
        else if (c.ecn_s == E_ECE_RCVD && !ece)
            c.ecn_s = E_IDLE;

        bool dupack = ack == old && ack < c.high;
        uint32_t old_dup = c.dupacks;
        if (dupack)
        {
            uint8_t cg = c.cong;
            if (cg != C_LOSS)
            {
                if (cg != C_RECOVERY)
                {
                    c.dupacks = old_dup + 1;
                    old_dup = c.dupacks - 1;
                }
                if (cg == C_OPEN)
                    c.cong = cg = C_DISORDER;
                if (cg == C_RECOVERY)
                    tb_add_reno_sack(c);
                else if (cg == C_DISORDER)
                {
                    if (c.dupacks == (uint32_t)DUPTHRESH && (old >= c.recover || !c.recover_active))
                        enter_recovery(x, f, now);
                    else if (tb_head_lost(c))
                        enter_recovery(x, f, now);
                }
            }
        }
        if (ack == old && ack == c.high)
        {
        }
        else if (ack == old && ack > c.high)
            c.nxt = ack;
        else if (ack == old)
            pkts_acked(x, c, 1, ack, now);
        else if (advancing)
        {
            uint32_t segs = delivered / MSS;
            c.ack_rem += delivered % MSS;
            if (c.ack_rem >= MSS)
            {
                segs++;
                c.ack_rem -= MSS;
            }
            if (!dupack)
                c.dupacks = 0;
            uint8_t cg = c.cong;
            if (ack < c.recover && cg == C_RECOVERY)
            {
                tb_mark_head_lost(c);
                if (!tb_retx_ends_at(c, ack + MSS))
                    do_retransmit(x, f, now);
                pkts_acked(x, c, 1, ack, now);
                new_ack(x, f, ack, c.first_partial, now);
                c.first_partial = 0;
            }
            else if (ack < c.recover && cg == C_LOSS)
            {
                pkts_acked(x, c, segs, ack, now);
                increase_window(x, c, segs, ack, now);
                new_ack(x, f, ack, true, now);
            }
            else if (cg == C_CWR)
            {
                pkts_acked(x, c, segs, ack, now);
                new_ack(x, f, ack, true, now);
            }
            else
            {
                bool exited = false;
                if (cg == C_OPEN)
                    pkts_acked(x, c, segs, ack, now);
                else if (cg == C_DISORDER)
                {
                    if (segs >= old_dup)
                        pkts_acked(x, c, segs - old_dup, ack, now);
                    if (!dupack)
                        c.cong = C_OPEN;
                }
                else if (cg == C_RECOVERY)
                {
                    c.first_partial = 1;
                    segs = (ack - old) / MSS;
                    pkts_acked(x, c, segs, ack, now);
                    c.cong = C_OPEN;
                    exited = true;
                    c.dupacks = 0;
                }
                else if (cg == C_LOSS)
                {
                    c.first_partial = 1;
                    segs = (ack - c.recover) / MSS;
                    pkts_acked(x, c, segs, ack, now);
                    c.cong = C_OPEN;
                }
                if (ack >= c.recover)
                    c.recover_active = 0;
                if (exited)
                {
                    new_ack(x, f, ack, true, now);
                    c.cwnd = c.ssthresh;
                }
                if (c.cong == C_OPEN)
                {
                    increase_window(x, c, segs, ack, now);
                    new_ack(x, f, ack, true, now);
                }
            }
        }
        if ((c.fin & FIN_SENT) && c.it_n == 0 && ack == c.high + 1)
            c.fin |= FIN_ACKED;
        if (advancing)
        {
            messages_acked(f, now);
            if (c.st == S_DONE)
                return;
        }
        send_pending(x, f, now);
    }
    // the retransmission timeout (src/step.cuh:1694-1725)
    void on_rto(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        c.rto_at = -1;
        if (c.st != S_EST)
            return;
        if (c.una >= c.tail)
        {
            if ((c.fin & (FIN_SENT | FIN_ACKED)) == FIN_SENT)
                ev[14]++;
            return;
        }
        ev[3]++;
        uint32_t before = inflight(c);
        c.dupacks = 0;
        tb_all_lost(c);
        c.recover = c.high;
        c.recover_active = 1;
        int64_t r = 2 * c.rto;
        c.rto = r < 60000000000000ll ? r : 60000000000000ll;
        c.h_n = 0;
        c.ssthresh = ssthresh_for(x, c, before);
        if (x.cubic)
            cub_loss(c, now);
        c.cong = C_LOSS;
        c.cwnd = MSS;
        send_pending(x, f, now);
    }

    // ------------------------------------------------------------ the receiving socket
    // (src/step.cuh:1727-2106)
    void send_synack(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        int64_t rto = CONN_TIMEOUT_PS << c.srv_syn_n;
        c.srv_syn_n++;
        if (c.r_synack_t < 0)
            c.r_synack_t = now;
        emit(x, f, F_SYN | F_ACK, 0, 0, SYN_FRAME, true, now);
        if (c.srv_syn_at < 0)
        {
            c.srv_syn_at = now + rto;
            timer_chain(x, c.dchain, rto);
        }
    }
    static void rcv_established(Conn &c)
    {
        if (!c.r_est)
        {
            c.r_est = 1;
            c.srv_syn_at = -1;
        }
    }
    void ce_state(HostCtx &x, int f, bool ce_now, int64_t now)
    {
        Conn &c = F[f];
        if (ce_now)
        {
            if (!c.ce_state && c.reserved && c.prior_flag)
                send_ack(x, f, c.prior_rcv_nxt, false, true, now);
            c.ce_state = 1;
            c.ecn_r = E_CE_RCVD;
        }
        else
        {
            if (c.ce_state && c.reserved && c.prior_flag)
                send_ack(x, f, c.prior_rcv_nxt, true, false, now);
            c.ce_state = 0;
            if (c.ecn_r == E_CE_RCVD || c.ecn_r == E_SENDING_ECE)
                c.ecn_r = E_IDLE;
        }
        c.prior_flag = 1;
        c.prior_rcv_nxt = c.rcv_nxt;
    }
    void rx_ecn(HostCtx &x, int f, const Frame &p, int64_t now)
    {
        if (!x.dctcp)
            return;
        Conn &c = F[f];
        bool ce = p.flags & F_CE;
        if (ce && c.ecn_ce_seq < (int64_t)p.sa)
        {
            c.ecn_ce_seq = p.sa;
            c.ecn_r = E_CE_RCVD;
            ce_state(x, f, true, now);
        }
        else if (p.flags & F_ECT)
            ce_state(x, f, false, now);
    }
    static void snd_ecn(HostCtx &x, Conn &c, const Frame &p)
    {
        if (!x.dctcp)
            return;
        if ((p.flags & F_CE) && !(c.s_ce & 1))
        {
            c.s_ce = 3;
            c.ecn_s = E_CE_RCVD;
        }
        else if (p.flags & F_ECT)
        {
            c.s_ce &= ~2;
            if (c.ecn_s == E_CE_RCVD || c.ecn_s == E_SENDING_ECE)
                c.ecn_s = E_IDLE;
        }
    }
    // TcpRxBuffer::Add for [seq, end): keep the out-of-order blocks disjoint and sorted
    static bool rx_add(Conn &c, uint32_t seq, uint32_t end_)
    {
        uint32_t head = std::max(seq, c.rcv_nxt);
        if (head >= end_)
            return false;
        const uint32_t cap = c.oo_cap, h0 = c.ooo_h;
        uint32_t n = c.ooo_n;
        auto at_ = [&](uint32_t i) { uint32_t v = h0 + i; return v >= cap ? v - cap : v; };
        std::vector<uint32_t> &bs = c.oo_s, &be = c.oo_e;
        uint32_t i;
        if (n == 0 || be[at_(n - 1)] < head)
            i = n;
        else
        {
            uint32_t lo = 0, hi = n - 1;
            while (lo < hi)
            {
                uint32_t m = (lo + hi) / 2;
                if (be[at_(m)] < head)
                    lo = m + 1;
                else
                    hi = m;
            }
            i = lo;
        }
        if (i < n && bs[at_(i)] <= head && be[at_(i)] >= end_)
            return false;
        uint32_t us = head, ue = end_, j = i;
        while (j < n && bs[at_(j)] <= ue)
        {
            us = std::min(us, bs[at_(j)]);
            ue = std::max(ue, be[at_(j)]);
            j++;
        }
        if (j == i)
        {
            if (n >= cap)
            {
                ev[6]++;
                return false;
            }
            for (uint32_t k = n; k > i; k--)
            {
                bs[at_(k)] = bs[at_(k - 1)];
                be[at_(k)] = be[at_(k - 1)];
            }
            n++;
        }
        else
        {
            uint32_t gone = j - i - 1;
            if (gone)
                for (uint32_t k = j; k < n; k++)
                {
                    bs[at_(k - gone)] = bs[at_(k)];
                    be[at_(k - gone)] = be[at_(k)];
                }
            n -= gone;
        }
        bs[at_(i)] = us;
        be[at_(i)] = ue;
        uint32_t h = h0;
        if (head == c.rcv_nxt)
        {
            c.rcv_nxt = be[at_(0)];
            h = at_(1);
            n--;
        }
        c.ooo_h = n ? h : 0;
        c.ooo_n = n;
        if (c.r_fin >= 0 && c.rcv_nxt == (uint32_t)c.r_fin)
            c.rcv_nxt++;
        return true;
    }
    static bool rx_finished(const Conn &c) { return c.r_fin >= 0 && (uint32_t)c.r_fin < c.rcv_nxt; }
    void send_fin_ack(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        c.delack_at = -1;
        c.delack_count = 0;
        emit(x, f, F_ACK | F_FIN, 0, c.rcv_nxt, CTRL_FRAME, true, now);
        if (c.r_frx_at < 0)
        {
            int64_t est = c.r_srtt >= 0 ? c.r_srtt : INIT_RTT_PS, v = 4 * c.r_rttvar;
            int64_t rto = std::max(est + std::max(v, CLOCK_GRAN_PS), MIN_RTO_PS);
            c.r_frx_at = now + rto;
            timer_chain(x, c.schain, rto);
        }
    }
    static int64_t last_rto(const Conn &c)
    {
        int64_t est = c.r_srtt >= 0 ? c.r_srtt : INIT_RTT_PS, v = 4 * c.r_rttvar;
        return est + std::max(v, CLOCK_GRAN_PS);
    }
    void do_peer_close(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        c.r_st = R_LAST_ACK;
        send_fin_ack(x, f, now);
        c.r_last_n = DATA_RETRIES;
        int64_t lr = last_rto(c);
        c.r_last_at = now + lr;
        timer_chain(x, c.dchain, lr);
    }
    static void rcv_closed(Conn &c)
    {
        c.r_st = R_CLOSED;
        c.r_last_at = -1;
        c.r_frx_at = -1;
        c.delack_at = -1;
    }
    void received_data(HostCtx &x, int f, uint32_t seq, uint32_t len, bool fin, int64_t now)
    {
        Conn &c = F[f];
        uint32_t expected = c.rcv_nxt;
        if (!rx_add(c, seq, seq + len))
        {
            send_ack(x, f, c.rcv_nxt, false, false, now);
            return;
        }
        if (expected < c.rcv_nxt && rx_finished(c) && !fin)
        {
            do_peer_close(x, f, now);
            return;
        }
        bool gap = c.ooo_n > 0, filled = c.rcv_nxt > expected + len;
        if (gap || filled)
        {
            c.reserved = 0;
            send_ack(x, f, c.rcv_nxt, false, false, now);
            return;
        }
        uint32_t cnt = c.delack_count + 1u;
        if (cnt >= DELACK_COUNT)
        {
            c.reserved = 0;
            send_ack(x, f, c.rcv_nxt, false, false, now);
        }
        else
        {
            c.delack_count = (uint8_t)cnt;
            c.reserved = 1;
            if (c.delack_at < 0)
            {
                c.delack_at = now + DELACK_PS;
                timer_chain(x, c.dchain, DELACK_PS);
            }
        }
    }
    void on_data(HostCtx &x, int f, const Frame &p, int64_t now)
    {
        Conn &c = F[f];
        if (!c.r_up || c.r_st == R_CLOSED)
            return;
        uint32_t seq = p.sa, len = payload(p), end_ = seq + len;
        bool fin = p.flags & F_FIN;
        if (len > 0 && c.r_est && (c.r_st == R_LAST_ACK ? seq != c.rcv_nxt : end_ < c.rcv_nxt))
        {
            send_ack(x, f, c.rcv_nxt, false, true, now);
            return;
        }
        rx_ecn(x, f, p, now);
        if ((p.flags & F_CWR) && c.r_est && c.r_st == R_OPEN && c.ecn_r != E_CE_RCVD)
            c.ecn_r = E_IDLE;
        if (!c.r_est)
        {
            if (fin && seq != c.rcv_nxt)
                return;
            rcv_established(c);
        }
        if (c.r_st == R_LAST_ACK)
        {
            if (fin || seq == c.rcv_nxt)
                rcv_closed(c);
            return;
        }
        if (!fin)
        {
            received_data(x, f, seq, len, false, now);
            return;
        }
        if (seq < c.rcv_nxt)
            return;
        c.r_fin = end_;
        if (c.rcv_nxt == end_)
            c.rcv_nxt++;
        if (len)
            received_data(x, f, seq, len, true, now);
        if (!rx_finished(c))
            return;
        do_peer_close(x, f, now);
    }
    void on_delack(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        c.delack_at = -1;
        c.delack_count = 0;
        c.reserved = 1;
        send_ack(x, f, c.rcv_nxt, false, false, now);
    }

    // ------------------------------------------------------------ application and dispatch
    static void add_active(std::vector<int> &list, int f)
    {
        if ((int)list.size() >= MAX_ACTIVE)
        {
            ev[5]++;
            return;
        }
        list.push_back(f);
    }
    // FlowSender::Write of every due message; Socket::Send defers SendPendingData by 1 ps
    // (src/step.cuh:2120-2145)
    void app_write(HostCtx &x, int f, int64_t now)
    {
        Conn &c = F[f];
        if (c.m_wr >= c.m_due)
            return;
        while (c.m_wr < c.m_due)
        {
            int m = c.m_first + c.m_wr;
            mg_est[m] = now;
            c.tail = mg_end[m];
            c.m_wr++;
        }
        if (c.fin_on && c.m_wr == c.m_cnt)
            c.fin |= FIN_CLOSE;
        uint32_t pipe = inflight(c);
        if (c.cwnd > pipe && c.syn_at < 0)
        {
            c.syn_at = now + 1;
            c.schain = chain_child(CODE_STEP, x.echain);
            due_insert(x, now + 1, f, 2);
        }
    }
    // one frame at its destination host (src/step.cuh:2147-2236)
    void on_frame(HostCtx &x, const Frame &p, int64_t now)
    {
        int f = (int)p.flow;
        Conn &c = F[f];
        if (p.to_sender)
        {
            if (p.flags & F_SYN)
            {
                if (c.st == S_SYN_SENT)
                {
                    c.st = S_EST;
                    c.syn_at = -1;
                    rtt_sample(c, now - c.timing_time);
                    emit(x, f, F_ACK, 0, 0, CTRL_FRAME, false, now);
                    x.echain = chain_child(CODE_NOW, x.echain);
                    app_write(x, f, now);
                }
            }
            else if (p.flags & F_FIN)
            {
                if (c.fin & FIN_TW)
                    return;
                snd_ecn(x, c, p);
                on_ack(x, f, p.sa, p.flags & F_ECE, now);
                c.fin |= FIN_TW;
                emit(x, f, F_ACK, 0, 1, CTRL_FRAME, false, now);
            }
            else
            {
                if (c.fin & FIN_TW)
                    return;
                snd_ecn(x, c, p);
                on_ack(x, f, p.sa, p.flags & F_ECE, now);
            }
            return;
        }
        if (!(p.flags & F_SYN) && c.r_up && c.r_srtt < 0 && c.r_synack_t >= 0)
        {
            int64_t m = now - c.r_synack_t;
            c.r_srtt = m;
            c.r_rttvar = m / 2;
        }
        if (p.flags & F_SYN)
        {
            if (!c.r_up)
            {
                c.r_up = 1;
                c.delack_count = DELACK_COUNT;
                add_active(x.host->act_rcv, f);
            }
            if (c.r_est)
                return;
            x.nic->woke = now;
            x.echain = chain_child(CODE_NOW, x.echain);
            send_synack(x, f, now);
        }
        else if (p.flags & (F_DATA | F_FIN))
            on_data(x, f, p, now);
        else if (c.r_up && c.r_st != R_CLOSED)
        {
            rx_ecn(x, f, p, now);
            rcv_established(c);
            if (c.r_st == R_LAST_ACK)
                rcv_closed(c);
        }
    }
    const Chain &due_chain(const Due &d) const
    {
        const Conn &c = F[d.flow];
        if (d.kind == 1 || d.kind == 3 || d.kind == 4)
            return c.dchain;
        if (d.kind == 5 || (d.kind == 2 && c.st == S_EST))
            return c.schain;
        return c.rchain;
    }
    // a timer and an arrival at one instant: the one ns-3 scheduled first runs first
    bool timer_before_arrival(const Due &d, const Frame &p)
    {
        Chain arr = chain_child(code_arrival(p.len), p.ch);
        int r = chain_cmp(due_chain(d), arr);
        if (r == 0)
            ev[25]++;
        return r <= 0;
    }

    // ------------------------------------------------------------ one slot of a host
    // (src/step.cuh:2263-2528): returns 1 if it sent or holds frames; *wake: its next slot
    int host_step(int h, int64_t t, std::vector<Due> &due, int64_t *wake)
    {
        const int64_t S0 = t * SLOT_PS, E0 = S0 + SLOT_PS;
        Host &H = hosts[h];
        int link = T.host_in[h];
        static const std::vector<Frame> none;
        const std::vector<Frame> *m2 = t >= 2 ? frames(t - 2, link) : nullptr, *m1 = t >= 1 ? frames(t - 1, link) : nullptr;
        if (!m2)
            m2 = &none;
        if (!m1)
            m1 = &none;
        int n2 = (int)m2->size(), n1 = (int)m1->size(), i2 = 0, i1 = 0;
        while (i2 < n2 && (*m2)[i2].ts < S0)
            i2++;
        bool input = i2 < n2 || (n1 > 0 && (*m1)[0].ts < E0);
        if (!input && t < H.wake && H.nic.size == 0)
        {
            *wake = H.wake;
            return 0;
        }
        Msg &out = message(t, h);
        out.tag = t;
        out.fr.clear();
        Fifo nic;
        nic.q = &H.nic;
        nic.slot_end = E0;
        nic.out = &out.fr;
        HostCtx x{this, &H, &nic, h, H.cc == CC_DCTCP, H.cc == CC_CUBIC, {}, &due, 0, 0, E0};
        // timers due in this slot, by time
        int nd = 0;
        int ns = (int)H.act_snd.size(), nr = (int)H.act_rcv.size();
        for (int i = 0; i < ns; i++)
        {
            int f = H.act_snd[i];
            int64_t a = F[f].rto_at, b = F[f].syn_at;
            if (a >= 0 && a < E0 && nd < MAX_DUE)
                due[nd++] = {a, f, 0};
            if (b >= 0 && b < E0 && nd < MAX_DUE)
                due[nd++] = {b, f, 2};
        }
        for (int i = 0; i < nr; i++)
        {
            int f = H.act_rcv[i];
            const Conn &c = F[f];
            int64_t a = c.delack_at, b = c.srv_syn_at, l = c.r_last_at, xx = c.r_frx_at;
            if (a >= 0 && a < E0 && nd < MAX_DUE)
                due[nd++] = {a, f, 1};
            if (b >= 0 && b < E0 && nd < MAX_DUE)
                due[nd++] = {b, f, 3};
            if (l >= 0 && l < E0 && nd < MAX_DUE)
                due[nd++] = {l, f, 4};
            if (xx >= 0 && xx < E0 && nd < MAX_DUE)
                due[nd++] = {xx, f, 5};
        }
        if (nd >= MAX_DUE)
            ev[7]++;
        for (int i = 1; i < nd; i++)
        {
            Due v = due[i];
            int j = i - 1;
            while (j >= 0 && due[j].at > v.at)
            {
                due[j + 1] = due[j];
                j--;
            }
            due[j + 1] = v;
        }
        x.nd = nd;
        x.id = 0;
        int pn = H.pend_next, pe = H.pend_end;
        for (;;)
        {
            int64_t tr = INT64_MAX, td = INT64_MAX, tp = INT64_MAX;
            if (i2 < n2)
                tr = (*m2)[i2].ts;
            else if (i1 < n1 && (*m1)[i1].ts < E0)
                tr = (*m1)[i1].ts;
            if (x.id < x.nd)
                td = due[x.id].at;
            if (pn < pe && mg_start[pend_msg[pn]] < E0)
                tp = mg_start[pend_msg[pn]];
            int64_t now = std::min(tr, std::min(td, tp));
            if (now == INT64_MAX)
                break;
            if (now < S0)
                now = S0;
            nic.advance(now);
            if (now == tp && F[mg_conn[pend_msg[pn]]].m_first != pend_msg[pn])
            {
                // a later message of a connection (FlowSender::SendMessage)
                int m = pend_msg[pn++], f = mg_conn[m];
                chain_root_msg(x.echain, m);
                F[f].m_due++;
                if (F[f].st == S_EST)
                    app_write(x, f, now);
            }
            else if (now == tp)
            {
                // a connection's application starts: SYN
                int m = pend_msg[pn++], f = mg_conn[m];
                Conn &c = F[f];
                chain_root(x.echain, h, f);
                c.m_due = 1;
                add_active(H.act_snd, f);
                c.st = S_SYN_SENT;
                if (x.cubic)
                    cub_hystart_reset(c, now);
                c.rto = CONN_TIMEOUT_PS;
                c.syn_at = now + CONN_TIMEOUT_PS;
                timer_chain(x, c.rchain, CONN_TIMEOUT_PS);
                c.timing_time = now;
                emit(x, f, F_SYN, 0, 0, SYN_FRAME, false, now);
            }
            else if (now == td && (now != tr || timer_before_arrival(due[x.id], i2 < n2 ? (*m2)[i2] : (*m1)[i1])))
            {
                Due d = due[x.id++];
                x.echain = due_chain(d);
                Conn &c = F[d.flow];
                if (d.kind == 0 && c.rto_at == d.at)
                    on_rto(x, d.flow, now);
                else if (d.kind == 1 && c.delack_at == d.at)
                    on_delack(x, d.flow, now);
                else if (d.kind == 4 && c.r_last_at == d.at && c.r_st == R_LAST_ACK)
                {
                    if (c.r_last_n == 0)
                        rcv_closed(c);
                    else
                    {
                        c.r_last_n--;
                        send_fin_ack(x, d.flow, now);
                        int64_t lr = last_rto(c);
                        c.r_last_at = now + lr;
                        timer_chain(x, c.dchain, lr);
                    }
                }
                else if (d.kind == 5 && c.r_frx_at == d.at && c.r_st == R_LAST_ACK)
                {
                    c.r_frx_at = -1;
                    send_fin_ack(x, d.flow, now);
                }
                else if (d.kind == 3 && c.srv_syn_at == d.at && !c.r_est)
                {
                    c.srv_syn_at = -1;
                    send_synack(x, d.flow, now);
                }
                else if (d.kind == 2 && c.syn_at == d.at && c.st == S_EST)
                {
                    c.syn_at = -1;
                    send_pending(x, d.flow, now);
                }
                else if (d.kind == 2 && c.syn_at == d.at && c.st == S_SYN_SENT)
                {
                    int tries = ++c.syn_tries;
                    c.rto = CONN_TIMEOUT_PS << tries;
                    c.syn_at = now + c.rto;
                    timer_chain(x, c.rchain, c.rto);
                    emit(x, d.flow, F_SYN, 0, 0, SYN_FRAME, false, now);
                }
            }
            else
            {
                Frame p = i2 < n2 ? (*m2)[i2++] : (*m1)[i1++];
                x.echain = chain_child(code_arrival(p.len), p.ch);
                on_frame(x, p, now);
            }
        }
        nic.advance(E0 - 1);
        H.pend_next = pn;
        // drop finished connections from the active lists; the next wake-up
        int64_t next = INT64_MAX;
        {
            int k = 0;
            for (size_t i = 0; i < H.act_snd.size(); i++)
            {
                int f = H.act_snd[i];
                if (F[f].st == S_DONE)
                    continue;
                H.act_snd[k++] = f;
                int64_t a = F[f].rto_at, b = F[f].syn_at;
                if (a >= 0 && a < next)
                    next = a;
                if (b >= 0 && b < next)
                    next = b;
            }
            H.act_snd.resize(k);
        }
        {
            int k = 0;
            for (size_t i = 0; i < H.act_rcv.size(); i++)
            {
                int f = H.act_rcv[i];
                const Conn &c = F[f];
                if (c.st == S_DONE && c.delack_at < 0 && c.r_last_at < 0 && c.r_frx_at < 0)
                    continue;
                H.act_rcv[k++] = f;
                for (int64_t v : {c.delack_at, c.srv_syn_at, c.r_last_at, c.r_frx_at})
                    if (v >= 0 && v < next)
                        next = v;
            }
            H.act_rcv.resize(k);
        }
        if (pn < pe && mg_start[pend_msg[pn]] < next)
            next = mg_start[pend_msg[pn]];
        H.wake = H.nic.size > 0 ? t + 1 : (next == INT64_MAX ? INT64_MAX : next / SLOT_PS);
        *wake = H.wake;
        return (!out.fr.empty() || H.nic.size > 0) ? 1 : 0;
    }
    // ------------------------------------------------------------ routing (src/step.cuh:2541-2630)
    int epoch_of(int64_t ts) const
    {
        int e = 0;
        while (e < T.n_fail && ts >= T.fail_t[e])
            e++;
        return e;
    }
    size_t rt_index(int ep, int x, int e) const { return ((size_t)ep * T.S + x) * T.S + e; }
    // the output port at switch sw towards host dst; up: the index among the equal-cost routes
    int route_up(int sw, int dst, int up, int ep) const
    {
        if (T.kind == TOPO_GRAPH)
        {
            int e = T.host_sw[dst];
            if (e == sw)
                return T.host_port[dst];
            return T.rt_ports[T.rt_off[rt_index(ep, sw, e)] + up];
        }
        if (T.kind == TOPO_LEAFSPINE)
        {
            if (sw < T.leaves)
                return dst / T.hpl == sw ? dst % T.hpl : T.hpl + up;
            return dst / T.hpl;
        }
        const int hk = T.hk, E = T.k * hk;
        int dst_pod = dst / (hk * hk), dst_edge = (dst / hk) % hk;
        if (sw < E)
            return (dst_pod == sw / hk && dst_edge == sw % hk) ? dst % hk : hk + up;
        if (sw < 2 * E)
            return dst_pod == (sw - E) / hk ? dst_edge : hk + up;
        return dst_pod;
    }
    int route(int sw, const Frame &p, int ep) const
    {
        if (T.kind == TOPO_GRAPH)
        {
            int e = T.host_sw[p.dst];
            if (e == sw)
                return T.host_port[p.dst];
            const Conn &c = F[p.flow];
            uint32_t h = p.to_sender ? c.hash_ack : c.hash_data;
            return route_up(sw, p.dst, (int)(h % T.rt_n[rt_index(ep, sw, e)]), ep);
        }
        return route_up(sw, p.dst, p.up, ep);
    }
    bool port_down(int sw, int q, int ep) const { return T.n_fail && T.port_dn[((size_t)ep * T.S + sw) * T.k + q]; }
    // Ipv4GlobalRouting's number of equal-cost routes (0 for an attached host)
    int routes_to(int sw, int dst, int ep) const
    {
        if (T.kind == TOPO_GRAPH)
            return T.host_sw[dst] == sw ? 0 : T.rt_n[rt_index(ep, sw, T.host_sw[dst])];
        if (T.kind == TOPO_LEAFSPINE)
            return sw < T.leaves ? (dst / T.hpl == sw ? 0 : T.ecmp_n) : 1;
        const int hk = T.hk, E = T.k * hk;
        int dst_pod = dst / (hk * hk), dst_edge = (dst / hk) % hk;
        if (sw < E)
            return dst_pod == sw / hk && dst_edge == sw % hk ? 0 : T.ecmp_n;
        if (sw < 2 * E)
            return dst_pod == (sw - E) / hk ? 1 : T.ecmp_n;
        return 1;
    }
    // simultaneous arrivals at a node: the longer frame first, then by the upstream starts'
    // chains, then (undecided) by flow id (src/step.cuh:2691-2704)
    static bool frame_before(const Frame &a, const Frame &b)
    {
        if (a.len != b.len)
            return a.len > b.len;
        int c = chain_cmp(a.ch, b.ch);
        if (c == 0)
            ev[25]++;
        return c != 0 ? c < 0 : a.flow < b.flow;
    }

    // ------------------------------------------------------------ one slot of a switch
    // (src/step.cuh:2706-2886): every ingress window, per-packet ECMP draws in node arrival
    // order, each egress's arrivals sorted by (time, frame_before) and enqueued; returns 1 if a
    // port sent or holds frames
    struct In
    {
        const std::vector<Frame> *m2 = nullptr, *m1 = nullptr;
        int i2 = 0, n2 = 0, n1 = 0;
        int count() const { return (n2 - i2) + n1; }
        const Frame &at(int j) const { return j < n2 - i2 ? (*m2)[i2 + j] : (*m1)[j - (n2 - i2)]; }
    };
    int switch_step(int sw, int64_t t)
    {
        const int k = T.k;
        const int64_t S0 = t * SLOT_PS, E0 = S0 + SLOT_PS;
        std::vector<In> in(k);
        for (int q = 0; q < k; q++)
        {
            int link = T.sw_in[(size_t)sw * k + q];
            if (link < 0)
                continue;
            In &x = in[q];
            if (t >= 2 && (x.m2 = frames(t - 2, link)))
            {
                x.n2 = (int)x.m2->size();
                while (x.i2 < x.n2 && (*x.m2)[x.i2].ts < S0)
                    x.i2++;
            }
            if (t >= 1 && (x.m1 = frames(t - 1, link)))
                while (x.n1 < (int)x.m1->size() && (*x.m1)[x.n1].ts < E0)
                    x.n1++;
        }
        // per-packet ECMP: one draw per forwarded packet, in the node's arrival order
        std::vector<std::vector<uint8_t>> up(k);
        if (!sw_rng.empty())
        {
            std::vector<int> pos(k, 0);
            for (int q = 0; q < k; q++)
                up[q].assign(in[q].count(), 0);
            Mrg &g = sw_rng[sw];
            for (;;)
            {
                int best = -1;
                const Frame *bf = nullptr;
                for (int q = 0; q < k; q++)
                {
                    if (pos[q] >= in[q].count())
                        continue;
                    const Frame &f = in[q].at(pos[q]);
                    if (best < 0 || f.ts < bf->ts || (f.ts == bf->ts && frame_before(f, *bf)))
                        best = q, bf = &f;
                }
                if (best < 0)
                    break;
                int ep = T.n_fail ? epoch_of(bf->ts) : 0;
                if (!port_down(sw, best, ep))
                {
                    int nr = routes_to(sw, bf->dst, ep);
                    up[best][pos[best]] = (uint8_t)(nr ? mrg_integer(g, (uint32_t)nr) : 0);
                }
                pos[best]++;
            }
        }
        // each arrival to its egress (ingress order, then arrival order)
        std::vector<std::vector<std::pair<int64_t, uint32_t>>> eg_in(k);
        for (int q = 0; q < k; q++)
            for (int j = 0; j < in[q].count(); j++)
            {
                const Frame &p = in[q].at(j);
                int ep = T.n_fail ? epoch_of(p.ts) : 0;
                if (port_down(sw, q, ep))
                {
                    ev[0]++;
                    continue;
                }
                int e = !sw_rng.empty() ? route_up(sw, p.dst, up[q][j], ep) : route(sw, p, ep);
                eg_in[e].push_back({p.ts, (uint32_t)(q << 16 | j)});
            }
        int active = 0;
        for (int e = 0; e < k; e++)
        {
            auto &v = eg_in[e];
            auto frame_at = [&](uint32_t r) -> const Frame & { return in[r >> 16].at(r & 0xFFFF); };
            for (size_t i = 1; i < v.size(); i++)
            {
                auto x = v[i];
                int j = (int)i - 1;
                while (j >= 0 && (v[j].first > x.first || (v[j].first == x.first && frame_before(frame_at(x.second), frame_at(v[j].second)))))
                {
                    v[j + 1] = v[j];
                    j--;
                }
                v[j + 1] = x;
            }
            int port = sw * k + e, out_link = T.H + port;
            Msg &out = message(t, out_link);
            out.tag = t;
            out.fr.clear();
            Fifo eg;
            eg.q = &ports[port];
            eg.slot_end = E0;
            eg.out = &out.fr;
            for (auto &r : v)
            {
                Frame p = frame_at(r.second);
                p.ch = chain_child(code_arrival(p.len), p.ch);
                eg.enqueue(p, p.ts);
            }
            eg.advance(E0 - 1);
            if (!out.fr.empty() || ports[port].size > 0)
                active = 1;
        }
        return active;
    }

    // ------------------------------------------------------------ the slot loop
    // (src/kernel.cuh:5-113): every host, every switch, then the ticket: finish, or the next
    // slot, skipping idle ones exactly (nothing sent in this slot or the previous one and nothing
    // queued: the next slot is the earliest host wake-up)
    int64_t run(double *seconds)
    {
        std::vector<Due> due(MAX_DUE);
        int64_t t = 0;
        unsigned prev = 1, cur = 1;
        auto t0 = std::chrono::steady_clock::now();
        if (M == 0)
        {
            *seconds = 0;
            return 0;
        }
        for (;;)
        {
            unsigned long long nw = ~0ull;
            for (int h = 0; h < T.H; h++)
            {
                int64_t w = INT64_MAX;
                cur += host_step(h, t, due, &w);
                nw = std::min(nw, (unsigned long long)w);
            }
            for (int sw = 0; sw < T.S; sw++)
                cur += switch_step(sw, t);
            bool finished = flows_done >= (unsigned)M || t + 1 >= max_slots;
            int64_t next = t + 1;
            if (cur == 0 && prev == 0 && nw != ~0ull && (int64_t)nw > next)
                next = (int64_t)nw < max_slots ? (int64_t)nw : max_slots;
            prev = cur;
            cur = 0;
            t = next;
            if (finished)
                break;
        }
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return t;
    }
};

// ---------------------------------------------------------------- set-up (src/main.cu:103-1015)
struct FlowIn
{
    int src, dst;
    int64_t size, start_ns, conn;
};

static std::vector<FlowIn> load_flows(const std::string &path, int H)
{
    std::vector<FlowIn> out;
    std::ifstream in(path);
    if (!in)
    {
        fprintf(stderr, "cannot open %s\n", path.c_str());
        exit(1);
    }
    std::string line;
    while (std::getline(in, line))
    {
        if (line.empty() || !isdigit((unsigned char)line[0]))
            continue;
        std::stringstream ss(line);
        std::string f;
        std::vector<int64_t> v;
        while (std::getline(ss, f, ','))
            v.push_back(std::stoll(f));
        if (v.size() == 3)
            out.push_back({(int)v[0], (int)((v[0] + H / 2) % H), v[1], v[2], -1});
        else if (v.size() == 4)
            out.push_back({(int)v[0], (int)v[1], v[2], v[3], -1});
        else if (v.size() == 5)
            out.push_back({(int)v[0], (int)v[1], v[2], v[3], v[4]});
    }
    return out;
}

static int die(const char *m)
{
    fprintf(stderr, "%s\n", m);
    exit(1);
}

} // namespace orc

int main(int argc, char **argv)
{
    using namespace orc;
    int k = 4, ecn_k = RED_K, queue = QDISC_CAP, fin = 0, spray = 0, leaves = 16, spines = 8, hpl = 16, hash_cols = 0;
    int rtt_ring = 65536, ooo_segs = 512;
    long long seed = 1, run = 1;
    int64_t max_slots = 2000000;
    std::string flows_file, out_file = "fct.csv", tcp = "dctcp", topo = "fattree", topo_file, fail_spec, ecmp_from, qtrace_file;
    for (int i = 1; i < argc; i++)
    {
        std::string a = argv[i];
        auto val = [&](const char *p) { return a.substr(strlen(p)); };
        auto is = [&](const char *p) { return a.rfind(p, 0) == 0; };
        if (is("--k="))
            k = std::stoi(val("--k="));
        else if (is("--flows="))
            flows_file = val("--flows=");
        else if (is("--out="))
            out_file = val("--out=");
        else if (is("--tcp="))
            tcp = val("--tcp=");
        else if (is("--qtrace="))
            qtrace_file = val("--qtrace="), qt_on = true;
        else if (is("--qtrace_from="))
            qt_from = std::stoll(val("--qtrace_from="));
        else if (is("--qtrace_until="))
            qt_until = std::stoll(val("--qtrace_until="));
        else if (is("--max_slots="))
            max_slots = std::stoll(val("--max_slots="));
        else if (is("--ecnK="))
            ecn_k = std::stoi(val("--ecnK="));
        else if (is("--queue="))
            queue = std::stoi(val("--queue="));
        else if (is("--topo="))
            topo = val("--topo=");
        else if (is("--topofile="))
            topo_file = val("--topofile=");
        else if (is("--fail="))
            fail_spec = val("--fail=");
        else if (is("--leaves="))
            leaves = std::stoi(val("--leaves="));
        else if (is("--spines="))
            spines = std::stoi(val("--spines="));
        else if (is("--hpl="))
            hpl = std::stoi(val("--hpl="));
        else if (is("--fin="))
            fin = std::stoi(val("--fin="));
        else if (is("--spray="))
            spray = std::stoi(val("--spray="));
        else if (is("--RngSeed="))
            seed = std::stoll(val("--RngSeed="));
        else if (is("--RngRun="))
            run = std::stoll(val("--RngRun="));
        else if (is("--chain_runs="))
            RUNS = std::stoi(val("--chain_runs="));
        else if (is("--hash_cols="))
            hash_cols = std::stoi(val("--hash_cols="));
        else if (is("--ecmp_from="))
            ecmp_from = val("--ecmp_from=");
        else if (is("--rtt_ring="))
            rtt_ring = std::stoi(val("--rtt_ring="));
        else if (is("--ooo_segs="))
            ooo_segs = std::stoi(val("--ooo_segs="));
        else if (a == "--general")
            ;
        else
        {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 1;
        }
    }
    if (RUNS != 9 && RUNS != 25)
        die("--chain_runs: 9 or 25");
    int cc = tcp == "dctcp" ? CC_DCTCP : tcp == "newreno" ? CC_NEWRENO : tcp == "cubic" ? CC_CUBIC : -1;
    if (flows_file.empty() || cc < 0)
        die("usage: tcp_oracle (--k=K | --topo=leafspine ... | --topo=graph --topofile=F) --flows=F [--out=F] [--tcp=...]");
    Sim *sp = new Sim;
    Sim &s = *sp;
    Topo &T = s.T;
    s.max_slots = max_slots;
    // ---- topology
    std::vector<std::pair<int, int>> links; // topology file: endpoints, host h -> h, switch j -> -1 - j
    std::vector<std::vector<std::pair<int, int>>> adj; // switch -> (port, neighbour switch)
    if (topo == "graph")
    {
        std::ifstream in(topo_file);
        if (!in)
            die("cannot open the topology file");
        int nh = 0, ns = 0;
        std::string key;
        auto node = [](const std::string &x) { return x[0] == 'h' ? std::stoi(x.substr(1)) : -1 - std::stoi(x.substr(1)); };
        while (in >> key)
        {
            if (key == "hosts")
                in >> nh;
            else if (key == "switches")
                in >> ns;
            else if (key == "link")
            {
                std::string a, b;
                in >> a >> b;
                links.push_back({node(a), node(b)});
            }
            else
                std::getline(in, key);
        }
        T.kind = TOPO_GRAPH;
        T.H = nh;
        T.S = ns;
        T.ecmp_n = 1;
        std::vector<int> deg(ns, 0);
        T.host_sw.assign(nh, -1);
        T.host_port.assign(nh, 0);
        adj.assign(ns, {});
        for (auto &l : links)
        {
            if (l.second >= 0 || (l.first >= 0 && T.host_sw[l.first] >= 0))
                die("hosts need exactly one link, as its first endpoint");
            int b = -1 - l.second, pb = deg[b]++;
            if (l.first >= 0)
            {
                T.host_sw[l.first] = b;
                T.host_port[l.first] = pb;
            }
            else
            {
                int a = -1 - l.first, pa = deg[a]++;
                adj[a].push_back({pa, b});
                adj[b].push_back({pb, a});
            }
        }
        k = 0;
        for (int d : deg)
            k = std::max(k, d);
        // failures T:sA:sB[:up] in time order; epoch i follows the i-th
        struct Fail
        {
            int64_t t;
            int a, pa, b, pb;
            bool up;
        };
        std::vector<Fail> fails;
        std::stringstream fs(fail_spec);
        for (std::string item; std::getline(fs, item, ',');)
        {
            if (item.empty())
                continue;
            std::vector<std::string> v;
            std::stringstream is(item);
            for (std::string x; std::getline(is, x, ':');)
                v.push_back(x);
            if (v.size() < 3)
                die("--fail: T:sA:sB[:up]");
            Fail f{std::stoll(v[0]) * 1000, std::stoi(v[1].substr(1)), -1, std::stoi(v[2].substr(1)), -1, v.size() > 3 && v[3] == "up"};
            for (auto &pn : adj[f.a])
                if (pn.second == f.b && f.pa < 0)
                    f.pa = pn.first;
            for (auto &pn : adj[f.b])
                if (pn.second == f.a && f.pb < 0)
                    f.pb = pn.first;
            if (f.pa < 0)
                die("--fail: no such link");
            fails.push_back(f);
        }
        std::stable_sort(fails.begin(), fails.end(), [](const Fail &x, const Fail &y) { return x.t < y.t; });
        int NE = (int)fails.size() + 1;
        T.n_fail = NE - 1;
        for (auto &f : fails)
            T.fail_t.push_back(f.t);
        T.rt_off.assign((size_t)NE * ns * ns, 0);
        T.rt_n.assign((size_t)NE * ns * ns, 0);
        T.port_dn.assign((size_t)NE * ns * k, 0);
        std::vector<uint8_t> dn((size_t)ns * k, 0);
        for (int ep = 0; ep < NE; ep++)
        {
            if (ep > 0)
            {
                const Fail &f = fails[ep - 1];
                dn[(size_t)f.a * k + f.pa] = dn[(size_t)f.b * k + f.pb] = f.up ? 0 : 1;
            }
            std::copy(dn.begin(), dn.end(), T.port_dn.begin() + (size_t)ep * ns * k);
            for (int e = 0; e < ns; e++)
            {
                // hop distances to switch e; from x, the ports towards a neighbour one hop
                // closer, in port order
                std::vector<int> dist(ns, -1), todo{e};
                dist[e] = 0;
                for (size_t i = 0; i < todo.size(); i++)
                    for (auto &pn : adj[todo[i]])
                        if (!dn[(size_t)todo[i] * k + pn.first] && dist[pn.second] < 0)
                        {
                            dist[pn.second] = dist[todo[i]] + 1;
                            todo.push_back(pn.second);
                        }
                for (int x = 0; x < ns; x++)
                {
                    if (x == e)
                        continue;
                    if (dist[x] < 0)
                        die("a switch cannot reach another");
                    std::vector<int> ps;
                    for (auto &pn : adj[x])
                        if (!dn[(size_t)x * k + pn.first] && dist[pn.second] == dist[x] - 1)
                            ps.push_back(pn.first);
                    std::sort(ps.begin(), ps.end());
                    size_t r = ((size_t)ep * ns + x) * ns + e;
                    T.rt_off[r] = (uint32_t)T.rt_ports.size();
                    T.rt_n[r] = (uint8_t)ps.size();
                    for (int p : ps)
                        T.rt_ports.push_back((uint8_t)p);
                }
            }
        }
    }
    else if (topo == "leafspine")
    {
        if (!fail_spec.empty())
            die("--fail needs --topo=graph");
        T.kind = TOPO_LEAFSPINE;
        k = std::max(hpl + spines, leaves);
        T.leaves = leaves, T.spines = spines, T.hpl = hpl;
        T.H = leaves * hpl;
        T.S = leaves + spines;
        T.ecmp_n = spines;
    }
    else
    {
        if (!fail_spec.empty())
            die("--fail needs --topo=graph");
        if (k < 4 || k > 32 || (k & (k - 1)))
            die("--k: a power of two in 4..32");
        T.kind = TOPO_FATTREE;
        T.H = k * k * k / 4;
        T.S = k * k + k * k / 4;
        T.ecmp_n = k / 2;
    }
    T.k = k;
    T.hk = k / 2;
    const int H = T.H, S = T.S, hk = T.hk, E = k * hk;
    s.L = H + S * k;
    // ---- messages over connections (src/main.cu:531-648)
    std::vector<FlowIn> flows = load_flows(flows_file, H);
    const int M = (int)flows.size();
    std::vector<int> msg_conn(M);
    std::vector<std::vector<int>> conn_msgs;
    {
        std::map<int64_t, int> by_id;
        for (int m = 0; m < M; m++)
        {
            const FlowIn &f = flows[m];
            if (f.src < 0 || f.src >= H || f.dst < 0 || f.dst >= H)
                die("a flow names a host outside the topology");
            int c;
            if (f.conn < 0 || !by_id.count(f.conn))
            {
                c = (int)conn_msgs.size();
                conn_msgs.push_back({});
                if (f.conn >= 0)
                    by_id[f.conn] = c;
            }
            else
            {
                c = by_id[f.conn];
                const FlowIn &first = flows[conn_msgs[c][0]];
                if (first.src != f.src || first.dst != f.dst || f.start_ns < flows[conn_msgs[c].back()].start_ns)
                    die("the flows of a connection need the same hosts and later starts");
            }
            conn_msgs[c].push_back(m);
            msg_conn[m] = c;
        }
    }
    const int NC = (int)conn_msgs.size();
    s.M = M;
    std::vector<int> mnew(M);
    {
        int j = 0;
        for (int c = 0; c < NC; c++)
            for (int m : conn_msgs[c])
                mnew[m] = j++;
    }
    auto is_first = [&](int i) { return conn_msgs[msg_conn[i]][0] == i; };
    std::vector<int> byold(M);
    for (int i = 0; i < M; i++)
        byold[i] = i;
    std::stable_sort(byold.begin(), byold.end(), [&](int a, int b) {
        if (flows[a].src != flows[b].src)
            return flows[a].src < flows[b].src;
        if (flows[a].start_ns != flows[b].start_ns)
            return flows[a].start_ns < flows[b].start_ns;
        bool fa = is_first(a), fb = is_first(b);
        if (fa != fb)
            return !fa;
        return fa ? msg_conn[a] < msg_conn[b] : a < b;
    });
    s.pend_msg.resize(M);
    for (int i = 0; i < M; i++)
        s.pend_msg[i] = mnew[byold[i]];
    s.hosts.resize(H);
    {
        int i = 0;
        for (int h = 0; h < H; h++)
        {
            s.hosts[h].pend_next = i;
            while (i < M && flows[byold[i]].src == h)
                i++;
            s.hosts[h].pend_end = i;
        }
    }
    s.mg_conn.resize(M);
    s.mg_size.resize(M);
    s.mg_end.resize(M);
    s.mg_start.resize(M);
    s.mg_est.assign(M, -1);
    s.mg_done.assign(M, -1);
    std::vector<uint64_t> conn_bytes(NC, 0);
    for (int c = 0; c < NC; c++)
    {
        uint64_t off = 0;
        for (int m : conn_msgs[c])
        {
            off += flows[m].size;
            if (off >= 0xFFFFFFFFull)
                die("a connection carries 4 GB or more");
            int j = mnew[m];
            s.mg_conn[j] = c;
            s.mg_size[j] = (uint32_t)flows[m].size;
            s.mg_end[j] = (uint32_t)off;
            s.mg_start[j] = flows[m].start_ns * 1000;
        }
        conn_bytes[c] = off;
    }
    // ---- connections, ECMP hashes (src/main.cu:649-769)
    s.F.resize(NC);
    for (int c = 0; c < NC; c++)
    {
        Conn &r = s.F[c];
        const FlowIn &first = flows[conn_msgs[c][0]];
        r.src = first.src;
        r.dst = first.dst;
        r.m_first = mnew[conn_msgs[c][0]];
        r.m_cnt = (int)conn_msgs[c].size();
        r.fin_on = fin != 0;
    }
    {
        std::vector<int> hsub(H);
        for (int h = 0; h < H; h++)
            hsub[h] = h;
        if (T.kind == TOPO_GRAPH)
            for (int i = 0; i < (int)links.size(); i++)
                if (links[i].first >= 0)
                    hsub[links[i].first] = i;
        auto ip = [&](int h) { return 0x0a000001u | (uint32_t)(hsub[h] & 0xFFFF) << 8; };
        auto key_hash = [](uint32_t a, uint32_t b, uint16_t pa, uint16_t pb) {
            uint8_t key[17] = {(uint8_t)(a >> 24), (uint8_t)(a >> 16), (uint8_t)(a >> 8), (uint8_t)a, (uint8_t)(b >> 24), (uint8_t)(b >> 16),
                               (uint8_t)(b >> 8), (uint8_t)b, 6, (uint8_t)(pa >> 8), (uint8_t)pa, (uint8_t)(pb >> 8), (uint8_t)pb};
            return hash32(key, 17);
        };
        std::vector<int> byhost(NC);
        for (int c = 0; c < NC; c++)
            byhost[c] = c;
        std::stable_sort(byhost.begin(), byhost.end(), [&](int a, int b) {
            const FlowIn &x = flows[conn_msgs[a][0]], &y = flows[conn_msgs[b][0]];
            return x.src != y.src ? x.src < y.src : x.start_ns < y.start_ns;
        });
        for (int i = 0, n = 0; i < NC; i++, n++)
        {
            int c = byhost[i];
            if (i > 0 && s.F[byhost[i - 1]].src != s.F[c].src)
                n = 0;
            if (n > 65535 - 49153 && ecmp_from.empty())
                die("a host binds too many connections: give --ecmp_from");
            uint16_t sport = (uint16_t)(49153 + n), dport = 1000;
            uint32_t a = ip(s.F[c].src), b = ip(s.F[c].dst);
            s.F[c].hash_data = key_hash(a, b, sport, dport);
            s.F[c].hash_ack = key_hash(b, a, dport, sport);
        }
        if (!ecmp_from.empty())
        {
            std::ifstream in(ecmp_from);
            std::string line;
            std::getline(in, line);
            int i = 0;
            while (i < M && std::getline(in, line))
            {
                std::vector<std::string> v;
                std::stringstream ss(line);
                for (std::string f; std::getline(ss, f, ',');)
                    v.push_back(f);
                if (v.size() < 8)
                    die("--ecmp_from: no hash columns");
                if (is_first(i))
                {
                    s.F[msg_conn[i]].hash_data = (uint32_t)std::stoul(v[6]);
                    s.F[msg_conn[i]].hash_ack = (uint32_t)std::stoul(v[7]);
                }
                i++;
            }
        }
    }
    // sent list / RTT history and out-of-order block capacities (src/main.cu:770-806)
    for (int c = 0; c < NC; c++)
    {
        Conn &r = s.F[c];
        uint64_t segs = (conn_bytes[c] + MSS - 1) / MSS + r.m_cnt + 1;
        r.it_cap = (uint32_t)std::min<uint64_t>(segs, (uint64_t)rtt_ring);
        r.oo_cap = (uint32_t)std::max(std::min<uint64_t>(segs / 2 + 2, (uint64_t)ooo_segs), std::min<uint64_t>((segs + 31) / 32, 1024));
        r.it_start.assign(r.it_cap, 0);
        r.it_fl.assign(r.it_cap, 0);
        r.h_seq.assign(r.it_cap, 0);
        r.h_cnt.assign(r.it_cap, 0);
        r.h_t.assign(r.it_cap, 0);
        r.oo_s.assign(r.oo_cap, 0);
        r.oo_e.assign(r.oo_cap, 0);
    }
    // ---- queues and links (src/main.cu:826-963)
    const int ring = queue + 2;
    for (int h = 0; h < H; h++)
    {
        Host &x = s.hosts[h];
        x.cc = (uint8_t)cc;
        x.nic.id = h;
        x.nic.ring.resize(ring);
        x.nic.red_k = ecn_k;
        x.nic.qcap = queue;
    }
    s.ports.resize((size_t)S * k);
    for (auto &p : s.ports)
    {
        p.id = H + (int)(&p - s.ports.data());
        p.ring.resize(ring);
        p.red_k = ecn_k;
        p.qcap = queue;
    }
    auto sw_link = [&](int sw, int p) { return H + sw * k + p; };
    T.host_in.assign(H, -1);
    T.sw_in.assign((size_t)S * k, -1);
    if (T.kind == TOPO_GRAPH)
    {
        for (int h = 0; h < H; h++)
        {
            T.host_in[h] = sw_link(T.host_sw[h], T.host_port[h]);
            T.sw_in[(size_t)T.host_sw[h] * k + T.host_port[h]] = h;
        }
        for (int x = 0; x < S; x++)
            for (auto &pn : adj[x])
            {
                int y = pn.second, nth = 0;
                for (auto &q : adj[x])
                    if (q.second == y && q.first < pn.first)
                        nth++;
                for (auto &q : adj[y])
                    if (q.second == x && nth-- == 0)
                    {
                        T.sw_in[(size_t)x * k + pn.first] = sw_link(y, q.first);
                        break;
                    }
            }
    }
    else if (T.kind == TOPO_LEAFSPINE)
    {
        for (int h = 0; h < H; h++)
        {
            T.host_in[h] = sw_link(h / hpl, h % hpl);
            T.sw_in[(size_t)(h / hpl) * k + h % hpl] = h;
        }
        for (int l = 0; l < leaves; l++)
            for (int j = 0; j < spines; j++)
            {
                int sp = leaves + j;
                T.sw_in[(size_t)l * k + hpl + j] = sw_link(sp, l);
                T.sw_in[(size_t)sp * k + l] = sw_link(l, hpl + j);
            }
    }
    else
    {
        for (int h = 0; h < H; h++)
        {
            T.host_in[h] = sw_link(h / hk, h % hk);
            T.sw_in[(size_t)(h / hk) * k + h % hk] = h;
        }
        for (int pod = 0; pod < k; pod++)
            for (int j = 0; j < hk; j++)
            {
                int agg = E + pod * hk + j;
                for (int x = 0; x < hk; x++)
                {
                    int e2 = pod * hk + x;
                    T.sw_in[(size_t)agg * k + x] = sw_link(e2, j + hk);
                    T.sw_in[(size_t)e2 * k + j + hk] = sw_link(agg, x);
                }
            }
        for (int i = 0; i < hk; i++)
            for (int j = 0; j < hk; j++)
            {
                int core = k * k + i * hk + j;
                for (int pod = 0; pod < k; pod++)
                {
                    int agg = E + pod * hk + i;
                    T.sw_in[(size_t)core * k + pod] = sw_link(agg, j + hk);
                    T.sw_in[(size_t)agg * k + j + hk] = sw_link(core, pod);
                }
            }
    }
    s.msg.resize((size_t)4 * s.L);
    if (spray)
        for (int sw = 0; sw < S; sw++)
            s.sw_rng.push_back(mrg_stream((uint32_t)seed, (uint64_t)(H + sw), (uint64_t)run));
    // ---- run
    double sec = 0;
    int64_t slots = s.run(&sec);
    printf("Simulation Duration: %.6fs\n", sec);
    printf("Executed Slots: %lld\n", (long long)slots);
    printf("flows %d completed %u wall %.4f s drops %llu marks %llu retx %llu rto %llu | tx-cap %llu active-ovf %llu ooo-ovf %llu due-ovf %llu"
           " | retx by state open %llu disorder %llu cwr %llu recovery %llu loss %llu | max-inflight %llu segs rtt-ovf %llu ring-ovf %llu"
           " | undecided-ties %llu\n",
           M, s.flows_done, sec, ev[0], ev[1], ev[2], ev[3], ev[4], ev[5], ev[6], ev[7], ev[8], ev[9], ev[10], ev[11], ev[12], ev[13], ev[14],
           ev[19], ev[25]);
    printf("packet-events %llu\n", tx_frames);
    if (ev[5] || ev[6] || ev[7] || ev[14] || ev[19])
        fprintf(stderr, "warning: a capacity overflowed (active/ooo/due/rtt/ring-ovf); the results may differ from ns-3\n");
    if (ev[25])
        fprintf(stderr, "warning: %llu ties between simultaneous events were undecided by their recorded histories (%d runs)\n", ev[25], RUNS);
    if (qt_on)
    {
        // src/main.cu:1077-1100: stable by (time, node, device), ns-3's node and device numbers
        auto node = [&](int q) { return q < H ? q : H + (q - H) / k; };
        auto dev = [&](int q) { return q < H ? 1 : (q - H) % k + 1; };
        std::stable_sort(qtrace.begin(), qtrace.end(), [&](const QRec &a, const QRec &b) {
            if (a.ts != b.ts)
                return a.ts < b.ts;
            if (node(a.queue) != node(b.queue))
                return node(a.queue) < node(b.queue);
            return dev(a.queue) < dev(b.queue);
        });
        FILE *q = fopen(qtrace_file.c_str(), "w");
        if (!q)
            die("cannot write the queue trace");
        for (const QRec &r : qtrace)
        {
            const Conn &c = s.F[r.flow];
            const char *kind = ((r.flags & F_SYN) && (r.flags & F_ACK)) ? "SA" : (r.flags & F_DATA) ? "D" : (r.flags & F_SYN) ? "S" : "A";
            fprintf(q, "%lld %d %d %d %d %s %u %d\n", (long long)r.ts, node(r.queue), dev(r.queue), r.to_sender ? c.dst : c.src,
                    r.to_sender ? c.src : c.dst, kind, r.sa, r.qlen);
        }
        fclose(q);
        printf("qtrace: %llu records\n", (unsigned long long)qtrace.size());
    }
    FILE *o = fopen(out_file.c_str(), "w");
    if (!o)
        die("cannot write the output");
    fprintf(o, "src,dst,flow_size,start_ns,established_ns,completed_ns%s\n", hash_cols ? ",hash_data,hash_ack" : "");
    for (int i = 0; i < M; i++)
    {
        int j = mnew[i];
        const Conn &c = s.F[msg_conn[i]];
        fprintf(o, "%d,%d,%lld,%lld,%lld,%lld", flows[i].src, flows[i].dst, (long long)flows[i].size, (long long)flows[i].start_ns,
                (long long)(s.mg_est[j] >= 0 ? s.mg_est[j] / 1000 : -1), (long long)(s.mg_done[j] >= 0 ? s.mg_done[j] / 1000 : -1));
        if (hash_cols)
            fprintf(o, ",%u,%u", c.hash_data, c.hash_ack);
        fprintf(o, "\n");
    }
    fclose(o);
    return 0;
}
