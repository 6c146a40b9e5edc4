// The TCP packet model on gfx950: one slot of a host (a lane runs its serial event loop) and one
// slot of a switch (a group of lanes of one wavefront, one lane per port), over the batch
// layout of tcp_state.h. The semantics follow TwiNS's model exactly (reference: src/step.cuh,
// src/kernel.cuh), so every flow's times equal sequential ns-3's; the comments cite the
// reference lines each part follows.
//
// Build with -ffp-contract=off: the only fused multiply-adds are the explicit fma() calls.
#pragma once
#include <cassert>
#include <cstddef>
#include <hip/hip_runtime.h>
#include "exact_math.h"
#include "tcp_state.h"

namespace tw
{

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int RUNS>
struct Model
{
    typedef BatchT<RUNS> Batch;
    typedef FrameT<RUNS> Frame;
    static constexpr int CW = RUNS + 1; // a chain's words: RUNS runs (code | count << 16), the root
    struct Chain
    {
        uint32_t r[CW];
    };

    // -------------------------------------------------------------- event-order chains
    // (src/step.cuh:58-159) the leads from an event back to its root, run-length encoded; codes
    // grow with the lead: ScheduleNow 1, Send's deferral 2, TransmitComplete len, arrival
    // 2048 + len, timers 4096 + lead/16 ns (saturating)
    static __device__ __forceinline__ uint32_t code_lead(int64_t lead_ps)
    {
        int64_t c = 4096 + lead_ps / 16000;
        return (uint32_t)(c < 0xFFFE ? c : 0xFFFE);
    }
    static __device__ __forceinline__ void chain_load(Chain &c, const uint32_t *p)
    {
#pragma unroll
        for (int i = 0; i < CW; i++)
            c.r[i] = p[i];
    }
    static __device__ __forceinline__ void chain_store(uint32_t *p, const Chain &c)
    {
#pragma unroll
        for (int i = 0; i < CW; i++)
            p[i] = c.r[i];
    }
    // an application start scheduled at set-up: host then connection (world-local)
    static __device__ __forceinline__ void chain_root(Chain &c, int host_local, int conn_local)
    {
#pragma unroll
        for (int i = 0; i < RUNS; i++)
            c.r[i] = 0;
        c.r[RUNS] = (uint32_t)(0x8000 | (host_local + 1)) | (uint32_t)(conn_local & 0xFFFF) << 16;
    }
    // a message write scheduled in main(), by message (world-local)
    static __device__ __forceinline__ void chain_root_msg(Chain &c, int msg_local)
    {
#pragma unroll
        for (int i = 0; i < RUNS; i++)
            c.r[i] = 0;
        c.r[RUNS] = (uint32_t)((1 + (msg_local >> 16)) & 0xFFFF) | (uint32_t)(msg_local & 0xFFFF) << 16;
    }
    // dst = [code] + parent (dst may alias parent): a repeated code extends the first run
    static __device__ __forceinline__ void chain_child(Chain &dst, uint32_t code, const Chain &parent)
    {
        const Chain p = parent;
        if ((p.r[0] & 0xFFFF) == code)
        {
            dst = p;
            if ((p.r[0] >> 16) < 0xFFFF)
                dst.r[0] += 1u << 16;
            return;
        }
        dst.r[0] = code | 1u << 16;
#pragma unroll
        for (int i = 1; i < RUNS; i++)
            dst.r[i] = p.r[i - 1];
        dst.r[RUNS] = p.r[RUNS];
    }
    // < 0: a runs first, > 0: b, 0: undecided. Larger codes (earlier parents) first; a complete
    // chain (fewer than RUNS runs) ends at its root, before any other parent; else the roots.
    // (a(i), b(i): the chains' words; chain_cmp reads plain memory, frames of link messages are
    // read device-coherently, word by word as the comparison needs them)
    template <typename GA, typename GB>
    static __device__ int chain_cmp_by(GA a, GB b)
    {
        int ia = 0, ib = 0;
        uint32_t ra = a(0) >> 16, rb = b(0) >> 16;
        for (;;)
        {
            uint32_t ca = a(ia) & 0xFFFF, cb = b(ib) & 0xFFFF;
            if (ca == 0 && cb != 0 && (a(RUNS) & 0xFFFF) != 0)
                return -1;
            if (cb == 0 && ca != 0 && (b(RUNS) & 0xFFFF) != 0)
                return 1;
            if (ca == 0 || cb == 0)
                break;
            if (ca != cb)
                return ca > cb ? -1 : 1;
            uint32_t m = ra < rb ? ra : rb;
            ra -= m;
            rb -= m;
            if (ra == 0)
            {
                if (++ia == RUNS)
                    break;
                ra = a(ia) >> 16;
            }
            if (rb == 0)
            {
                if (++ib == RUNS)
                    break;
                rb = b(ib) >> 16;
            }
        }
        uint32_t xa = a(RUNS) << 16 | a(RUNS) >> 16, xb = b(RUNS) << 16 | b(RUNS) >> 16; // the root keys
        if (xa == 0 || xb == 0 || xa == xb)
            return 0;
        return xa < xb ? -1 : 1;
    }
    static __device__ int chain_cmp(const uint32_t *a, const uint32_t *b)
    {
        return chain_cmp_by([&](int i) { return a[i]; }, [&](int i) { return b[i]; });
    }
    // a link message's frame, read device-coherently: its head (all but the chain), its chain
    struct FrameHead
    {
        int64_t ts;
        uint32_t flow, sa;
        uint16_t len, dst;
        uint8_t flags, to_sender, up, pad;
    };
    static_assert(sizeof(FrameHead) == 24 && offsetof(Frame, chain) == 24, "a frame's head");
    static __device__ __forceinline__ FrameHead msg_head(const Frame *p) { return coh_get(reinterpret_cast<const FrameHead *>(p)); }
    static __device__ __forceinline__ void msg_chain(Chain &c, const Frame *p)
    {
#pragma unroll
        for (int i = 0; i < CW; i++)
            c.r[i] = coh_load(&p->chain[i]);
    }
    static __device__ __forceinline__ void count(WorldRun *R, int i, unsigned long long n = 1) { atomicAdd(&R->ev[i], n); }

    // -------------------------------------------------------------- egress queue
    // One sending queue per link (a host's NIC, a switch port): frames wait in a ring until the
    // link is free; a frame whose transmission starts in this slot goes to the link's message for
    // the slot with its arrival time at the far end. The queue's state lives in registers for the
    // entity's slot (load / store). Admission is RED with MinTh = MaxTh = K and QW = 1 on the
    // queue-disc length ns-3 would see (src/step.cuh:167-317 restates the same rules).
    struct Fifo
    {
        const Batch *B;
        WorldRun *R;
        int lid;
        Frame *ring;
        int cap, head, size, red_k, limit, sent, early;
        int64_t link_free, last_start, woke, idle_start, slot_end;
        int64_t head_ts; // the first queued frame's enqueue time (while size > 0)
        Chain last; // the chain of the transmission in progress (kept in registers for the slot)
        Frame *out;

        __device__ __forceinline__ void load(const Batch &b, WorldRun *r, int link, int64_t t, int ring_t)
        {
            B = &b;
            R = r;
            lid = link;
            ring = b.qring + b.q_base[link];
            cap = b.q_cap[link];
            head = b.q_head[link];
            size = b.q_size[link];
            red_k = b.q_red_k[link];
            limit = b.q_limit[link];
            link_free = b.q_link_free[link];
            last_start = b.q_last_start[link];
            chain_load(last, b.q_last_chain + (size_t)link * CW);
            woke = idle_start = -1;
            slot_end = (t + 1) * SLOT_PS;
            sent = early = 0;
            head_ts = size > 0 ? ring[head].ts : 0;
            out = b.msg + ((size_t)ring_t * b.nlinks + link) * MAX_TX;
        }
        // the slot's message count word (msg_word) and the queue back to memory
        __device__ __forceinline__ void store(int64_t t, int ring_t)
        {
            coh_store(B->msg_cnt + (size_t)ring_t * B->nlinks + lid, msg_word(t, early, sent));
            B->q_head[lid] = head;
            B->q_size[lid] = size;
            B->q_link_free[lid] = link_free;
            B->q_last_start[lid] = last_start;
            chain_store(B->q_last_chain + (size_t)lid * CW, last);
            if (sent)
                count(R, EV_PKT_EVENTS, (unsigned long long)sent);
        }
        // Starts, in order, the frames whose transmission begins by `until` within the slot. The
        // event that starts a frame is the previous frame's TransmitComplete when the frame had
        // to wait for the link (its chain: the last transmission's, one more run), else the
        // frame's own enqueue (a link that never sent counts as idle).
        __device__ void advance(int64_t until)
        {
            while (size > 0)
            {
                if (sent >= MAX_TX)
                {
                    count(R, EV_TXCAP);
                    return;
                }
                // (the head's time is kept in a register, so a call that starts nothing reads no frame)
                const int64_t start = head_ts > link_free ? head_ts : link_free;
                if (start >= slot_end || start > until)
                    return;
                Frame o = ring[head];
                if (head_ts <= link_free && link_free > last_start)
                {
                    chain_child(last, (uint32_t)((link_free - last_start) / PS_PER_BYTE), last);
                    chain_store(o.chain, last);
                }
                else
                {
                    chain_load(last, o.chain);
                    idle_start = start;
                }
                last_start = start;
                link_free = start + (int64_t)o.len * PS_PER_BYTE;
                o.ts = link_free + PROP_PS; // its arrival at the far end
                early += o.ts < slot_end + SLOT_PS;
                coh_put(out + sent, o); // (read by the link's far end in a later slot)
                sent++;
                head = head + 1 < cap ? head + 1 : 0;
                if (--size > 0)
                    head_ts = ring[head].ts;
            }
        }
        // a queue-trace record (src/step.cuh:266-279), stored field by field (no record on the stack)
        __device__ __forceinline__ void trace(int64_t now, int qlen, uint32_t flow, uint32_t sa, uint16_t len, uint8_t flags, uint8_t to_sender)
        {
            const unsigned long long i = atomicAdd(B->qt.cnt, 1ull);
            if (i < B->qt.cap)
            {
                QRec *r = B->qt.rec + i;
                r->ts = now;
                r->link = lid;
                r->flow = flow;
                r->sa = sa;
                r->len = len;
                r->qlen = (uint16_t)qlen;
                r->flags = flags;
                r->to_sender = to_sender;
            }
        }
        // The queue-disc length ns-3's RED sees at `now`. The first queued frame sits in the
        // device's one-frame queue (outside the queue disc) once the device has pulled it: when a
        // transmission started before `now`, or when the device's Wake ran at this very instant
        // on an idle link; until then it still counts.
        __device__ __forceinline__ int qdisc_len(int64_t now) const
        {
            const bool pulled = last_start < now || (woke == now && idle_start == now);
            return size > 0 && link_free > now && pulled ? size - 1 : size;
        }
        // a frame offered at `now`: RED marks (ECN-capable) or drops it at K, the queue limit
        // drops it; false if dropped
        __device__ bool enqueue(const Frame &f, int64_t now)
        {
            advance(now);
            const int len = qdisc_len(now);
            Frame g = f;
            g.ts = now;
            const bool over_k = len >= red_k && len > 1;
            if (over_k && !(g.flags & F_ECT))
            {
                count(R, EV_DROP);
                return false;
            }
            if (over_k)
            {
                g.flags |= F_CE;
                count(R, EV_MARK);
            }
            if (len >= limit)
            {
                count(R, EV_DROP);
                return false;
            }
            if (B->qt.rec && now >= B->qt.from && now < B->qt.until)
                trace(now, len + 1, g.flow, g.sa, g.len, g.flags, g.to_sender);
            if (size >= cap)
            {
                count(R, EV_RING_OVF);
                return false;
            }
            const int at = head + size < cap ? head + size : head + size - cap;
            ring[at] = g;
            if (size++ == 0)
                head_ts = g.ts;
            return true;
        }
    };

    // -------------------------------------------------------------- host context
    struct Host
    {
        const Batch *B;
        const WorldDesc *W;
        WorldRun *R;
        Fifo nic;
        int h, hl; // global, world-local
        bool dctcp, cubic;
        Chain echain; // the chain of the event being processed
        Due *due;
        int nd, id;
        int64_t due_head; // due[id].at (INT64_MAX when none is left): the next timer of the slot
        int64_t slot_end;
        // the records of the connection side an event works on, copied into this lane's LDS
        // slots for the event (snd_open / rcv_open) and written back once at its end
        SndState *ls;
        RcvState *lr;
        ConnStatic *lc; // the event's connection's static record, copied with its side's record
#ifndef __HIP_DEVICE_COMPILE__
        int lc_flow = -1; // (CPU builds: which connection lc holds, checked by conn())
#endif
    };
    // an event on connection f's sending (receiving) side: its record from device memory into
    // the lane's LDS slot, and back when the event ends (a record is 16 (8) independent 16-byte
    // loads, issued together, instead of a dependent load per field use)
    // (dst and src: one record type, in LDS and in device memory, either way round)
    template <typename D, typename S>
    static __device__ __forceinline__ void rec_move(D *dst, const S *src)
    {
        static_assert(sizeof(D) == sizeof(S) && sizeof(D) % 16 == 0, "records move in 16-byte pieces");
#pragma unroll
        for (int i = 0; i < (int)(sizeof(D) / 16); i++)
            reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
    }
    // the static record comes along with either side's record (its loads issued with the record's)
    static __device__ __forceinline__ void conn_open(Host &x, int f)
    {
        rec_move(x.lc, &x.B->cs[f]);
#ifndef __HIP_DEVICE_COMPILE__
        x.lc_flow = f;
#endif
    }
    // the open connection's static record (connection f's: an event works on one connection)
    static __device__ __forceinline__ const ConnStatic &conn(const Host &x, int f)
    {
#ifndef __HIP_DEVICE_COMPILE__
        assert(x.lc_flow == f);
#endif
        return *x.lc;
    }
    static __device__ __forceinline__ SndState &snd_open(Host &x, int f)
    {
        rec_move(x.ls, &x.B->snd[f]);
        conn_open(x, f);
        return *x.ls;
    }
    // The receiver's host reads a sender's end state (compact_rcv): the 4-byte word holding
    // SndState::st goes out device-coherently once the sender is done.
    static constexpr size_t ST_WORD = offsetof(SndState, st) & ~(size_t)3;
    static constexpr int ST_SHIFT = (int)(offsetof(SndState, st) & 3) * 8;
    static __device__ __forceinline__ void snd_close(Host &x, int f)
    {
        rec_move(&x.B->snd[f], x.ls);
        if (x.ls->st == S_DONE)
            coh_store(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(&x.B->snd[f]) + ST_WORD),
                      *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(x.ls) + ST_WORD));
    }
    static __device__ __forceinline__ bool snd_done_seen(const Batch &B, int f)
    {
        const uint32_t w = coh_load(reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(&B.snd[f]) + ST_WORD));
        return (uint8_t)(w >> ST_SHIFT) == S_DONE;
    }
    static __device__ __forceinline__ RcvState &rcv_open(Host &x, int f)
    {
        rec_move(x.lr, &x.B->rcv[f]);
        conn_open(x, f);
        return *x.lr;
    }
    static __device__ __forceinline__ void rcv_close(Host &x, int f) { rec_move(&x.B->rcv[f], x.lr); }

    // Active lists: a host's sending and receiving connections. A connection leaves its list at
    // the end of a slot in which it is finished (sender: closed; receiver: the sender closed and
    // no receiver timer pending), as in the reference (src/step.cuh:2470-2510; with a full list
    // the add is refused and counted, so the removal's timing decides results in that regime);
    // compact() drops those and returns the earliest timer of the rest.
    // (Both walk their list COMPACT_RUN entries at a time: the entries' records are all read
    // before any entry is written back, so the reads of a run are in flight together.)
    static constexpr int COMPACT_RUN = 4;
    static __device__ __forceinline__ int64_t compact_snd(const Batch &B, int h)
    {
        int64_t next = INT64_MAX;
        int32_t *l = B.act_snd + (size_t)h * MAX_ACTIVE;
        int k = 0, n = B.n_snd[h];
        for (int i0 = 0; i0 < n; i0 += COMPACT_RUN)
        {
            int f[COMPACT_RUN];
            bool done[COMPACT_RUN];
            int64_t rto[COMPACT_RUN], syn[COMPACT_RUN];
#pragma unroll
            for (int j = 0; j < COMPACT_RUN; j++)
                if (i0 + j < n)
                {
                    f[j] = l[i0 + j];
                    const SndState &s = B.snd[f[j]];
                    done[j] = s.st == S_DONE;
                    rto[j] = s.rto_at;
                    syn[j] = s.syn_at;
                }
#pragma unroll
            for (int j = 0; j < COMPACT_RUN; j++)
                if (i0 + j < n && !done[j])
                {
                    l[k++] = f[j];
                    if (rto[j] >= 0 && rto[j] < next)
                        next = rto[j];
                    if (syn[j] >= 0 && syn[j] < next)
                        next = syn[j];
                }
        }
        B.n_snd[h] = k;
        return next;
    }
    static __device__ __forceinline__ int64_t compact_rcv(const Batch &B, int h)
    {
        int64_t next = INT64_MAX;
        int32_t *l = B.act_rcv + (size_t)h * MAX_ACTIVE;
        int k = 0, n = B.n_rcv[h];
        for (int i0 = 0; i0 < n; i0 += COMPACT_RUN)
        {
            int f[COMPACT_RUN];
            bool gone[COMPACT_RUN];
            int64_t tmin[COMPACT_RUN];
#pragma unroll
            for (int j = 0; j < COMPACT_RUN; j++)
                if (i0 + j < n)
                {
                    f[j] = l[i0 + j];
                    const RcvState &r = B.rcv[f[j]];
                    const int64_t v[4] = {r.delack_at, r.srv_syn_at, r.r_last_at, r.r_frx_at};
                    gone[j] = snd_done_seen(B, f[j]) && v[0] < 0 && v[2] < 0 && v[3] < 0;
                    tmin[j] = INT64_MAX;
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (v[q] >= 0 && v[q] < tmin[j])
                            tmin[j] = v[q];
                }
#pragma unroll
            for (int j = 0; j < COMPACT_RUN; j++)
                if (i0 + j < n && !gone[j])
                {
                    l[k++] = f[j];
                    next = tmin[j] < next ? tmin[j] : next;
                }
        }
        B.n_rcv[h] = k;
        return next;
    }

    static __device__ void due_insert(Host &x, int64_t at, int f, int kind)
    {
        if (at >= x.slot_end)
            return;
        if (x.nd >= MAX_DUE)
        {
            count(x.R, EV_DUE_OVF);
            return;
        }
        int j = x.nd;
        while (j > x.id && x.due[j - 1].at > at)
        {
            x.due[j] = x.due[j - 1];
            j--;
        }
        x.due[j] = {at, f, kind};
        x.nd++;
        if (j == x.id)
            x.due_head = at;
    }
    static __device__ __forceinline__ void timer_chain(Host &x, uint32_t *base, int f, int64_t lead)
    {
        Chain c;
        chain_child(c, code_lead(lead), x.echain);
        chain_store(base + (size_t)f * CW, c);
    }
    static __device__ void emit(Host &x, int f, uint8_t flags, uint32_t seq, uint32_t ack, uint16_t len, bool to_sender, int64_t now)
    {
        const ConnStatic &cs = conn(x, f);
        Frame fr;
        fr.ts = now;
        fr.flow = (uint32_t)f;
        fr.sa = (flags & F_DATA) ? seq : ack;
        fr.len = len;
        fr.dst = (uint16_t)((to_sender ? cs.src : cs.dst) - x.W->host0);
        fr.flags = flags | (x.dctcp ? F_ECT : 0);
        fr.to_sender = to_sender ? 1 : 0;
        fr.up = (uint8_t)((to_sender ? cs.hash_ack : cs.hash_data) % (uint32_t)x.W->ecmp_n);
        fr.pad = 0;
        chain_store(fr.chain, x.echain);
        x.nic.enqueue(fr, now);
    }

    // -------------------------------------------------------------- the sent ledger
    // ns-3's TcpTxBuffer sent list as this model uses it (no SACK blocks; Reno's emulated SACK on
    // duplicate ACKs): one entry per run of transmitted bytes - its first byte and its marks
    // (IT_SACKED, IT_LOST, IT_RETX) - in a ring per connection, oldest first; an entry ends where
    // the next begins (the last at s.high). The connection's record keeps what decisions need
    // without walking the ring: bytes per mark, the sacked entries, and the retransmission
    // candidates (lost, neither retransmitted nor sacked) with a lower bound of the first.
    struct Ledger
    {
        SndState &s;
        uint32_t TW_G *first;
        uint8_t TW_G *marks;
        uint32_t cap;
        __device__ Ledger(const Batch &B, SndState &rec, const ConnStatic &c) : s(rec)
        {
            first = B.it_start + c.it_off;
            marks = B.it_fl + c.it_off;
            cap = c.it_cap;
        }
        __device__ uint32_t pos(uint32_t i) const { return s.it_head + i < cap ? s.it_head + i : s.it_head + i - cap; }
        __device__ uint32_t lo(uint32_t i) const { return first[pos(i)]; }
        __device__ uint32_t hi(uint32_t i) const { return i + 1 < s.it_n ? first[pos(i + 1)] : s.high; }
        __device__ uint8_t mk(uint32_t i) const { return marks[pos(i)]; }
    };
    static __device__ __forceinline__ bool lg_candidate(uint8_t m) { return (m & (IT_LOST | IT_RETX | IT_SACKED)) == IT_LOST; }
    // adds (sign 1) or removes (sign -1) an entry of `len` bytes with marks m from the totals
    static __device__ __forceinline__ void lg_count(SndState &s, uint8_t m, uint32_t len, int sign)
    {
        const uint32_t d = sign > 0 ? len : 0u - len, one = sign > 0 ? 1u : 0u - 1u;
        if (m & IT_SACKED)
        {
            s.sack_b += d;
            s.sacked_n += one;
        }
        if (m & IT_LOST)
            s.lost_b += d;
        if (m & IT_RETX)
            s.rtx_b += d;
        if (lg_candidate(m))
            s.cand_n += one;
    }
    // entry i's marks become m
    static __device__ void lg_remark(Ledger &L, uint32_t i, uint8_t m)
    {
        const uint32_t k = L.pos(i);
        const uint8_t was = L.marks[k];
        if (was == m)
            return;
        const uint32_t len = L.hi(i) - L.first[k];
        lg_count(L.s, was, len, -1);
        lg_count(L.s, m, len, 1);
        if (lg_candidate(m) && !lg_candidate(was) && i < L.s.cand)
            L.s.cand = i;
        L.marks[k] = m;
    }
    // a new entry at place i (the entries from i on move up one); its bytes were counted before
    // (a split) or are new (unmarked); false if the ring is full
    static __device__ bool lg_insert(Ledger &L, WorldRun *R, uint32_t i, uint32_t start, uint8_t m)
    {
        SndState &s = L.s;
        if (s.it_n >= L.cap)
        {
            count(R, EV_RTT_OVF);
            return false;
        }
        for (uint32_t j = s.it_n; j > i; j--)
        {
            const uint32_t to = L.pos(j), from = L.pos(j - 1);
            L.first[to] = L.first[from];
            L.marks[to] = L.marks[from];
        }
        L.first[L.pos(i)] = start;
        L.marks[L.pos(i)] = m;
        s.it_n++;
        if (s.cand > 0 && s.cand >= i)
            s.cand++;
        if (lg_candidate(m))
        {
            s.cand_n++;
            s.cand = i < s.cand ? i : s.cand;
        }
        return true;
    }
    // entry i leaves (the ones above move down); the caller has settled its counts
    static __device__ void lg_erase(Ledger &L, uint32_t i)
    {
        SndState &s = L.s;
        for (uint32_t j = i + 1; j < s.it_n; j++)
        {
            const uint32_t to = L.pos(j - 1), from = L.pos(j);
            L.first[to] = L.first[from];
            L.marks[to] = L.marks[from];
        }
        s.it_n--;
        if (s.cand > i)
            s.cand--;
    }
    // the first entry starting at or after seq
    static __device__ uint32_t lg_find(const Ledger &L, uint32_t seq)
    {
        uint32_t a = 0, b = L.s.it_n;
        while (a < b)
        {
            const uint32_t m = (a + b) / 2;
            if (L.lo(m) < seq)
                a = m + 1;
            else
                b = m;
        }
        return a;
    }
    static __device__ __forceinline__ uint32_t lg_sent(const SndState &s) { return s.it_n ? s.high - s.una : 0u; }
    // bytes in flight (TcpTxBuffer::BytesInFlight without SACK blocks)
    static __device__ __forceinline__ uint32_t inflight(const SndState &s) { return lg_sent(s) - (s.sack_b + s.lost_b) + s.rtx_b; }
    // a duplicate ACK in recovery: the next unsacked entry after the ones already sacked (Reno's
    // emulation of a SACK block)
    static __device__ void lg_dup_sack(Ledger &L)
    {
        SndState &s = L.s;
        uint32_t i = s.sacked_n + 1;
        for (; i < s.it_n; i++)
            if (!(L.mk(i) & IT_SACKED))
            {
                lg_remark(L, i, L.mk(i) | IT_SACKED);
                s.sack_seen = 1;
                s.hisack = L.lo(i);
                return;
            }
    }
    // the oldest entry lost (and no longer sacked or retransmitted)
    static __device__ void lg_head_lost(Ledger &L)
    {
        if (L.s.it_n)
            lg_remark(L, 0, (uint8_t)(IT_LOST | (L.mk(0) & ~(IT_SACKED | IT_RETX))));
    }
    // a retransmission timeout: every entry lost, nothing sacked or retransmitted
    static __device__ void lg_all_lost(Ledger &L)
    {
        SndState &s = L.s;
        for (uint32_t i = 0; i < s.it_n; i++)
            L.marks[L.pos(i)] = IT_LOST;
        s.lost_b = lg_sent(s);
        s.sack_b = s.rtx_b = 0;
        s.sacked_n = 0;
        s.cand_n = s.it_n;
        s.cand = 0;
        s.sack_seen = 0;
        s.hisack = 0;
    }
    // the bytes up to ack are acknowledged (DiscardUpTo): entries wholly below it leave, one cut by
    // it starts there; returns the bytes of the entries that left
    static __device__ uint32_t lg_release(Ledger &L, uint32_t ack)
    {
        SndState &s = L.s;
        uint32_t freed = 0, gone = 0;
        while (s.it_n)
        {
            const uint32_t k = L.pos(0), a = L.first[k], b = L.hi(0);
            if (b > ack)
            {
                if (ack > a)
                {
                    // the entry's acknowledged front leaves its mark totals
                    const uint8_t m = L.marks[k];
                    const uint32_t cut = ack - a;
                    s.sack_b -= m & IT_SACKED ? cut : 0u;
                    s.lost_b -= m & IT_LOST ? cut : 0u;
                    s.rtx_b -= m & IT_RETX ? cut : 0u;
                    L.first[k] = ack;
                }
                break;
            }
            if (L.marks[k])
                lg_remark(L, 0, 0);
            s.it_head = L.pos(1);
            s.it_n--;
            gone++;
            freed += b - a;
        }
        if (gone)
            s.cand = s.cand > gone ? s.cand - gone : 0;
        s.una = ack;
        if (s.it_n && (L.mk(0) & IT_SACKED))
        {
            lg_remark(L, 0, (uint8_t)(L.mk(0) & ~IT_SACKED));
            lg_dup_sack(L);
            lg_head_lost(L);
        }
        if (s.hisack <= ack)
            s.sack_seen = 0, s.hisack = 0;
        return freed;
    }
    // where the next segment starts: the first retransmission candidate below the highest sack,
    // else new data; false: nothing to send
    static __device__ bool lg_next(Ledger &L, uint32_t *seq)
    {
        SndState &s = L.s;
        for (uint32_t i = s.cand; s.cand_n && i < s.it_n; i++)
        {
            if (!lg_candidate(L.mk(i)))
                continue;
            s.cand = i;
            if (s.sack_seen && L.lo(i) >= s.hisack)
                break;
            *seq = L.lo(i);
            return true;
        }
        if (s.high >= s.tail)
            return false;
        *seq = s.high;
        return true;
    }
    // is the oldest entry lost below the highest sack (a fast retransmit due)?
    static __device__ bool lg_head_is_lost(const Ledger &L)
    {
        const SndState &s = L.s;
        return s.sack_seen && s.una < s.hisack && s.it_n && (L.mk(0) & IT_LOST);
    }
    // does a retransmitted, unsacked entry end exactly at x?
    static __device__ bool lg_retx_ends_at(const Ledger &L, uint32_t x)
    {
        const SndState &s = L.s;
        if (!s.it_n || x <= s.una || x > s.high)
            return false;
        const uint32_t n = lg_find(L, x);
        if (!n || L.hi(n - 1) != x)
            return false;
        const uint8_t m = L.mk(n - 1);
        return (m & IT_RETX) && !(m & IT_SACKED);
    }
    // num bytes sent again from entry i (GetTransmittedSegment): it merges with the next entry
    // when that one is unsacked and equally lost (their retransmitted marks settled first), or is
    // split at num; the result is marked retransmitted. Returns the bytes.
    static __device__ uint32_t lg_resend(Ledger &L, WorldRun *R, uint32_t i, uint32_t num)
    {
        SndState &s = L.s;
        uint32_t len = L.hi(i) - L.lo(i);
        const bool has_next = i + 1 < s.it_n;
        const uint8_t a = L.mk(i), b = has_next ? L.mk(i + 1) : 0;
        const bool mergeable = has_next && !(b & IT_SACKED) && (a & IT_LOST) == (b & IT_LOST);
        const uint32_t most = mergeable ? len + (L.hi(i + 1) - L.lo(i + 1)) : len;
        const uint32_t sz = num < most ? num : most;
        if (sz > len)
        {
            // entry i + 1 joins entry i: only one of them may stay retransmitted-marked
            uint8_t nb = b;
            if ((a ^ b) & IT_RETX)
            {
                if (a & IT_RETX)
                    lg_remark(L, i, (uint8_t)(a & ~IT_RETX));
                else
                    lg_remark(L, i + 1, nb = (uint8_t)(b & ~IT_RETX));
            }
            if (lg_candidate(nb))
                s.cand_n--;
            lg_erase(L, i + 1);
            len = L.hi(i) - L.lo(i);
        }
        if (sz < len)
            lg_insert(L, R, i + 1, L.lo(i) + sz, L.mk(i));
        if (!(L.mk(i) & IT_RETX))
            lg_remark(L, i, L.mk(i) | IT_RETX);
        return sz;
    }
    // the segment of up to num bytes at seq (CopyFromSequence): new data appends an entry, data
    // sent before is resent from its entry (*retx); a segment across s.high sends only the part
    // below it. 0: nothing (or the ledger cannot hold it)
    static __device__ uint32_t lg_take(Ledger &L, WorldRun *R, uint32_t num, uint32_t seq, bool *retx)
    {
        SndState &s = L.s;
        for (;;)
        {
            const uint32_t sz = s.tail <= seq ? 0u : num < s.tail - seq ? num : s.tail - seq;
            if (!sz)
                return 0;
            if (s.high <= seq)
            {
                if (!lg_insert(L, R, s.it_n, seq, 0))
                    return 0;
                s.high = seq + sz;
                *retx = false;
                return sz;
            }
            if (s.high >= seq + sz)
            {
                const uint32_t i = lg_find(L, seq);
                if (i >= s.it_n || L.lo(i) != seq)
                {
                    count(R, EV_RTT_OVF);
                    return 0;
                }
                *retx = true;
                return lg_resend(L, R, i, sz);
            }
            num = s.high - seq;
        }
    }

    // -------------------------------------------------------------- RTT
    // Each transmission is noted (first byte, length; the length's top bit once any of it is
    // resent, which removes it from sampling: Karn) in a ring beside the ledger; the oldest
    // unretransmitted record an ACK covers gives a sample (ns-3's RttMeanDeviation, in integer
    // picoseconds with its shifts), and the records it covers leave.
    static __device__ __forceinline__ uint32_t rtt_pos(const SndState &s, uint32_t cap, uint32_t j)
    {
        return s.h_head + j < cap ? s.h_head + j : s.h_head + j - cap;
    }
    static __device__ void rtt_note(Host &x, int f, uint32_t seq, uint32_t sz, bool retx, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        const ConnStatic &c = conn(x, f);
        uint32_t *hs = B.h_seq + c.it_off, *hc = B.h_cnt + c.it_off;
        if (retx)
        {
            // the record holding seq: resent up to seq + sz
            for (uint32_t j = 0; j < s.h_n; j++)
            {
                const uint32_t k = rtt_pos(s, c.it_cap, j);
                if (seq >= hs[k] && seq - hs[k] < (hc[k] & 0x7FFFFFFFu))
                {
                    hc[k] = 0x80000000u | (seq + sz - hs[k]);
                    return;
                }
            }
            return;
        }
        if (s.h_n >= c.it_cap)
        {
            count(x.R, EV_RTT_OVF);
            return;
        }
        const uint32_t k = rtt_pos(s, c.it_cap, s.h_n++);
        hs[k] = seq;
        hc[k] = sz;
        B.h_t[c.it_off + k] = now;
    }
    // a measurement m: the smoothed RTT and its variation, then the timeout (at least MinRto)
    static __device__ void rtt_sample(SndState &s, int64_t m)
    {
        if (s.srtt < 0)
        {
            s.srtt = m;
            s.rttvar = m / 2;
        }
        else
        {
            const int64_t err = m - s.srtt, mag = err < 0 ? -err : err;
            s.srtt = ((s.srtt << 3) + err) >> 3;
            s.rttvar = ((s.rttvar << 2) - s.rttvar + mag) >> 2;
        }
        const int64_t spread = 4 * s.rttvar > CLOCK_GRAN_PS ? 4 * s.rttvar : CLOCK_GRAN_PS;
        s.rto = s.srtt + spread > MIN_RTO_PS ? s.srtt + spread : MIN_RTO_PS;
    }
    static __device__ void rtt_on_ack(Host &x, int f, uint32_t ack, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        if (!s.h_n)
            return;
        const ConnStatic &c = conn(x, f);
        const uint32_t *hs = B.h_seq + c.it_off, *hc = B.h_cnt + c.it_off;
        const uint32_t k0 = rtt_pos(s, c.it_cap, 0);
        if (!(hc[k0] & 0x80000000u) && ack >= hs[k0] + hc[k0])
            rtt_sample(s, now - B.h_t[c.it_off + k0]);
        for (; s.h_n; s.h_n--)
        {
            const uint32_t k = rtt_pos(s, c.it_cap, 0);
            if (hs[k] + (hc[k] & 0x7FFFFFFFu) > ack)
                break;
            s.h_head = rtt_pos(s, c.it_cap, 1);
        }
    }

    // -------------------------------------------------------------- congestion control
    // Three controllers over the sender's record, each with ns-3's hooks: the slow-start threshold
    // after a loss or an echo (GetSsThresh), the per-ACK measurement (PktsAcked) and the window
    // growth (IncreaseWindow). Cubic follows ns-3's TcpCubic (with HyStart), DCTCP its alpha EWMA
    // per window of data, NewReno slow start and congestion avoidance.
    static __device__ __forceinline__ uint32_t to_u32(double d) { return (uint32_t)(uint64_t)(int64_t)d; }
    static __device__ void cubic_round_start(SndState &s, int64_t now)
    {
        s.cub_roundstart = now;
        s.cub_lastack = now;
        s.cub_endseq = s.st == S_EST ? s.high + 1 : 0;
        s.cub_currrtt = INT64_MIN;
        s.cub_samples = 0;
    }
    // ACKs per window increment in congestion avoidance (TcpCubic::Update): the cubic curve's
    // target at now + delay_min against the window, bounded by the TCP-friendly estimate
    static __device__ uint32_t cubic_acks_per_inc(SndState &s, uint32_t segs, int64_t now)
    {
        const double C = 0.4, BETA = 0.7;
        const uint32_t w = s.cwnd / MSS;
        s.cub_ackcnt += segs;
        if (s.cub_epoch == INT64_MIN)
        {
            // a new epoch: the curve's origin and the time K to reach it
            s.cub_epoch = now;
            s.cub_ackcnt = segs;
            s.cub_tcpcwnd = w;
            const bool below = s.cub_lastmax > w;
            s.cub_origin = below ? s.cub_lastmax : w;
            s.cub_k = below ? tw_pow((double)(s.cub_lastmax - w) / C, 1 / 3.) : 0.0;
        }
        const int64_t dt = (int64_t)((uint64_t)now + (uint64_t)s.cub_delaymin - (uint64_t)s.cub_epoch);
        const double t = tw_seconds(dt), k = s.cub_k;
        const uint32_t off = to_u32(C * tw_pow(t < k ? k - t : t - k, 3));
        const uint32_t target = t < k ? s.cub_origin - off : s.cub_origin + off;
        uint32_t n = target > w ? w / (target - w) : 100 * w;
        if (!s.cub_lastmax && n > 20)
            n = 20;
        // the TCP-friendly window: one segment per `per` ACKs
        const uint32_t per = (w * to_u32(8 * (1024 + BETA * 1024) / 3 / (1024 - BETA * 1024))) >> 3;
        for (; per > 0 && s.cub_ackcnt > per; s.cub_ackcnt -= per)
            s.cub_tcpcwnd++;
        if (s.cub_tcpcwnd > w)
        {
            const uint32_t cap = w / (s.cub_tcpcwnd - w);
            n = n > cap ? cap : n;
        }
        return n < 2 ? 2 : n;
    }
    // HyStart's delay check and the minimum delay (TcpCubic::PktsAcked)
    static __device__ void cubic_measure(SndState &s, int64_t now)
    {
        const int64_t ACK_DELTA = 2000000000, DELAY_MIN = 4000000000, DELAY_MAX = 1000000000000, CUBIC_DELTA = 10000000000;
        if (s.cub_epoch != INT64_MIN && now - s.cub_epoch < CUBIC_DELTA)
            return;
        const int64_t rtt = s.srtt;
        if (s.cub_delaymin == INT64_MIN || rtt < s.cub_delaymin)
            s.cub_delaymin = rtt;
        const bool slow_start_window = s.cwnd <= s.ssthresh && s.cwnd >= 16 * MSS;
        if (!slow_start_window || s.cub_found)
            return;
        if (now - s.cub_lastack <= ACK_DELTA)
        {
            // the ACK train
            s.cub_lastack = now;
            if (now - s.cub_roundstart > s.cub_delaymin)
                s.cub_found = 1;
        }
        if (s.cub_samples >= 8)
        {
            // the delay increase
            const int64_t dm = s.cub_delaymin, eta = dm < DELAY_MIN ? DELAY_MIN : dm > DELAY_MAX ? DELAY_MAX : dm;
            if (s.cub_currrtt > dm + eta)
                s.cub_found = 1;
        }
        else
        {
            if (s.cub_currrtt == INT64_MIN || rtt < s.cub_currrtt)
                s.cub_currrtt = rtt;
            s.cub_samples++;
        }
        if (s.cub_found)
            s.ssthresh = s.cwnd;
    }
    static __device__ void cubic_grow(SndState &s, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (!s.cwnd_limited)
            return;
        if (s.cwnd < s.ssthresh)
        {
            if (ack + 1 > s.cub_endseq)
                cubic_round_start(s, now);
            s.cwnd += segs * MSS;
            segs = 0;
        }
        if (segs == 0 || s.cwnd < s.ssthresh)
            return;
        s.cub_cnt += segs;
        const uint32_t n = cubic_acks_per_inc(s, segs, now);
        if (s.cub_cnt >= n)
        {
            s.cub_cnt -= n;
            s.cwnd += MSS;
        }
    }
    // the threshold after a loss: beta times the window (fast convergence lowers the remembered
    // maximum when the window is below it); a new epoch starts
    static __device__ uint32_t cubic_ssthresh(SndState &s)
    {
        const double BETA = 0.7;
        const uint32_t w = s.cwnd / MSS;
        s.cub_lastmax = w < s.cub_lastmax ? to_u32((w * (1 + BETA)) / 2) : w;
        s.cub_epoch = INT64_MIN;
        const uint32_t th = to_u32(w * BETA);
        return (th < 2 ? 2 : th) * MSS;
    }
    static __device__ void cubic_timeout(SndState &s, int64_t now)
    {
        s.cub_origin = 0;
        s.cub_k = 0;
        s.cub_ackcnt = 0;
        s.cub_tcpcwnd = 0;
        s.cub_delaymin = INT64_MIN;
        s.cub_found = 0;
        cubic_round_start(s, now);
    }
    // DCTCP's alpha: the marked share of the bytes acknowledged over each window of data
    static __device__ void dctcp_measure(SndState &s, uint32_t segs, uint32_t ack)
    {
        const uint32_t bytes = segs * MSS;
        s.acked_total += bytes;
        if (s.ecn_s == E_ECE_RCVD)
            s.acked_ecn += bytes;
        if (!s.next_seq_flag)
        {
            s.next_seq_flag = 1;
            s.next_seq = s.nxt;
        }
        if (ack < s.next_seq)
            return;
        const double share = s.acked_total ? (double)s.acked_ecn / s.acked_total : 0.0;
        s.alpha = __builtin_fma(1.0 - DCTCP_G, s.alpha, DCTCP_G * share);
        s.next_seq = s.nxt;
        s.acked_total = s.acked_ecn = 0;
    }
    // DCTCP's growth: slow start up to the threshold, then a segment per window's worth of ACKs
    static __device__ void dctcp_grow(SndState &s, uint32_t segs)
    {
        if (s.cwnd < s.ssthresh)
        {
            if (segs)
            {
                const uint64_t w = (uint64_t)s.cwnd + (uint64_t)segs * MSS;
                s.cwnd = (uint32_t)(w < s.ssthresh ? w : s.ssthresh);
            }
            return;
        }
        const uint32_t per = s.cwnd / MSS ? s.cwnd / MSS : 1;
        uint32_t acc = s.cwnd_cnt;
        if (acc >= per)
        {
            acc = 0;
            s.cwnd += MSS;
        }
        acc += segs;
        if (acc >= per)
        {
            s.cwnd += acc / per * MSS;
            acc %= per;
        }
        s.cwnd_cnt = acc;
    }
    // NewReno: one segment per ACK in slow start, then about MSS^2 / cwnd bytes (at least 1)
    static __device__ void newreno_grow(SndState &s, uint32_t segs)
    {
        if (s.cwnd < s.ssthresh && segs)
        {
            s.cwnd += MSS;
            segs--;
        }
        if (s.cwnd >= s.ssthresh && segs)
        {
            const double add = (double)MSS * MSS / s.cwnd;
            s.cwnd += (uint32_t)(add < 1.0 ? 1.0 : add);
        }
    }
    static __device__ uint32_t cc_ssthresh(Host &x, SndState &s, uint32_t in_flight)
    {
        if (x.cubic)
            return cubic_ssthresh(s);
        if (x.dctcp)
            return (uint32_t)((1.0 - s.alpha / 2.0) * s.cwnd);
        return in_flight / 2 > 2 * MSS ? in_flight / 2 : 2 * MSS;
    }
    static __device__ void cc_acked(Host &x, SndState &s, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (x.cubic)
            cubic_measure(s, now);
        else if (x.dctcp)
            dctcp_measure(s, segs, ack);
    }
    static __device__ void cc_grow(Host &x, SndState &s, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (x.cubic)
            cubic_grow(s, segs, ack, now);
        else if (x.dctcp)
            dctcp_grow(s, segs);
        else
            newreno_grow(s, segs);
    }

    // -------------------------------------------------------------- the sending endpoint
    // A connection's sending side (TcpSocketBase with its TxBuffer, as the model restates it):
    // segments leave through the host's queue, the retransmission timer follows the oldest
    // unacknowledged byte, and ACKs drive ns-3's congestion states (open, disorder, CWR,
    // recovery, loss). Every function works on the record of the event's connection in LDS.
    // one segment of up to max_size bytes at seq (SendDataPacket): CWR on the first new segment
    // after an echo, FIN on the last byte once the application closed; the timer is armed if
    // idle; returns its bytes
    static __device__ uint32_t tx_segment(Host &x, int f, uint32_t seq, uint32_t max_size, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        Ledger L(B, s, conn(x, f));
        bool retx = false;
        const uint32_t len = lg_take(L, x.R, max_size, seq, &retx);
        if (!len)
            return 0;
        uint8_t flags = F_DATA;
        if (!retx && s.ecn_s == E_ECE_RCVD && s.ecn_echo_seq > s.ecn_cwr_seq)
        {
            flags |= F_CWR;
            s.ecn_s = E_CWR_SENT;
            s.ecn_cwr_seq = seq;
        }
        if ((s.fin & FIN_CLOSE) && seq + len == s.tail)
        {
            flags |= F_FIN;
            s.fin |= FIN_SENT;
        }
        emit(x, f, flags, seq, 0, (uint16_t)(len + 42), false, now);
        if (retx)
        {
            count(x.R, EV_RETX);
            count(x.R, EV_RETX_BY_STATE + s.cong);
        }
        if (s.rto_at < 0)
        {
            s.rto_at = now + s.rto;
            timer_chain(x, B.rchain, f, s.rto);
        }
        s.cwnd_limited = s.cwnd < inflight(s) + MSS;
        rtt_note(x, f, seq, len, retx, now);
        s.high = seq + len > s.high ? seq + len : s.high;
        atomicMax(&x.R->ev[EV_MAX_INFLIGHT], (unsigned long long)s.it_n);
        return len;
    }
    // sends while the window has room (SendPendingData): whole segments only, unless the rest of
    // the data is smaller; nothing more once a FIN is out in the open state
    static __device__ void tx_fill(Host &x, int f, int64_t now)
    {
        SndState &s = *x.ls;
        if (s.st != S_EST)
            return;
        Ledger L(*x.B, s, conn(x, f));
        for (int n = 0; n < 4096; n++)
        {
            const bool fin_out = (s.fin & (FIN_SENT | FIN_ACKED)) == FIN_SENT;
            if (s.cong == C_OPEN && fin_out)
                return;
            const uint32_t pipe = inflight(s), room = s.cwnd > pipe ? s.cwnd - pipe : 0;
            uint32_t seq;
            if (!room || !lg_next(L, &seq))
                return;
            if (room < MSS && s.tail - seq > room)
                return; // (no runt segment while more data waits)
            s.nxt = seq;
            const uint32_t len = tx_segment(x, f, seq, room < MSS ? room : MSS, now);
            if (!len)
                return;
            s.nxt = seq + len;
        }
    }
    // one segment again from the first candidate (else the oldest byte)
    static __device__ void tx_resend_one(Host &x, int f, int64_t now)
    {
        SndState &s = *x.ls;
        Ledger L(*x.B, s, conn(x, f));
        uint32_t seq;
        if (!lg_next(L, &seq))
            seq = s.una;
        if (s.una >= s.tail)
            return;
        s.nxt = seq;
        tx_segment(x, f, seq, MSS, now);
    }
    // the retransmission timer after new data is acknowledged (NewAck): restarted from the RTT
    // estimate when asked, stopped once all data (and no FIN) is acknowledged
    static __device__ void rto_after_ack(Host &x, int f, uint32_t ack, bool restart, int64_t now)
    {
        SndState &s = *x.ls;
        if (restart)
        {
            const int64_t base = s.srtt >= 0 ? s.srtt : INIT_RTT_PS, spread = 4 * s.rttvar > CLOCK_GRAN_PS ? 4 * s.rttvar : CLOCK_GRAN_PS;
            s.rto = base + spread > MIN_RTO_PS ? base + spread : MIN_RTO_PS;
            s.rto_at = now + s.rto;
            timer_chain(x, x.B->rchain, f, s.rto);
        }
        if (ack >= s.tail && !(s.fin & FIN_SENT))
            s.rto_at = -1;
        s.nxt = ack > s.nxt ? ack : s.nxt;
    }
    // fast retransmit: the oldest segment lost, the window halved (or as the controller says),
    // the segment sent again
    static __device__ void fast_retransmit(Host &x, int f, int64_t now)
    {
        SndState &s = *x.ls;
        Ledger L(*x.B, s, conn(x, f));
        lg_dup_sack(L);
        lg_head_lost(L);
        s.cong = C_RECOVERY;
        s.recover = s.high;
        s.recover_active = 1;
        s.ssthresh = cc_ssthresh(x, s, inflight(s) + MSS);
        s.cwnd = s.ssthresh;
        tx_segment(x, f, s.una, MSS, now);
    }
    // messages wholly acknowledged complete (their completion time is the flow's); the connection
    // is done with its last message (and its FIN, if it sends one)
    static __device__ void msgs_complete(Host &x, int f, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        const ConnStatic &c = conn(x, f);
        for (; s.m_dn < s.m_wr; s.m_dn++)
        {
            const int m = c.m_first + s.m_dn;
            if (s.una < B.mg_end[m])
                break;
            B.mg_done[m] = now;
            atomicAdd(&x.R->flows_done, 1u);
        }
        if (s.m_dn == c.m_cnt && (!c.fin_on || (s.fin & FIN_ACKED)))
        {
            s.st = S_DONE;
            s.rto_at = -1;
        }
    }
    // DCTCP's response to an echo (EnterCwr): a new echo beyond the last one cuts the window once
    static __device__ void ack_echo(Host &x, SndState &s, uint32_t ack, bool ece, bool advancing)
    {
        if (!(advancing && x.dctcp && ece))
        {
            if (!ece && s.ecn_s == E_ECE_RCVD)
                s.ecn_s = E_IDLE;
            return;
        }
        if ((int64_t)ack <= s.ecn_echo_seq)
            return;
        s.ecn_echo_seq = ack;
        s.ecn_s = E_ECE_RCVD;
        if (s.cong == C_CWR)
            return;
        s.ssthresh = cc_ssthresh(x, s, inflight(s));
        s.cwnd = s.ssthresh;
        s.cong = C_CWR;
        s.recover = s.high;
    }
    // a duplicate ACK (below s.high): counted outside recovery and loss; in recovery it sacks
    // the next entry; the third (or a lost head below the highest sack) starts fast retransmit
    static __device__ void ack_duplicate(Host &x, int f, SndState &s, uint32_t una_before, int64_t now)
    {
        uint8_t state = s.cong;
        if (state == C_LOSS)
            return;
        if (state != C_RECOVERY)
            s.dupacks++;
        if (state == C_OPEN)
            s.cong = state = C_DISORDER;
        Ledger L(*x.B, s, conn(x, f));
        if (state == C_RECOVERY)
            lg_dup_sack(L);
        else if (state == C_DISORDER)
        {
            const bool third = s.dupacks == (uint32_t)DUPTHRESH && (!s.recover_active || una_before >= s.recover);
            if (third || lg_head_is_lost(L))
                fast_retransmit(x, f, now);
        }
    }
    // an ACK of new data below the recovery point: partial (recovery: the next hole is resent,
    // the timer restarted on the first partial ACK only) or during loss (slow start again)
    static __device__ void ack_partial(Host &x, int f, SndState &s, uint32_t ack, uint32_t segs, int64_t now)
    {
        if (s.cong == C_RECOVERY)
        {
            Ledger L(*x.B, s, conn(x, f));
            lg_head_lost(L);
            if (!lg_retx_ends_at(L, ack + MSS))
                tx_resend_one(x, f, now);
            cc_acked(x, s, 1, ack, now);
            rto_after_ack(x, f, ack, s.first_partial, now);
            s.first_partial = 0;
            return;
        }
        cc_acked(x, s, segs, ack, now);
        cc_grow(x, s, segs, ack, now);
        rto_after_ack(x, f, ack, true, now);
    }
    // an ACK of new data otherwise: disorder and the ends of recovery and loss return to the
    // open state (the window settles at the threshold after recovery); open grows the window
    static __device__ void ack_advance(Host &x, int f, SndState &s, uint32_t ack, uint32_t una_before, uint32_t segs, uint32_t dups, int64_t now)
    {
        const uint8_t state = s.cong;
        if (state == C_CWR)
        {
            cc_acked(x, s, segs, ack, now);
            rto_after_ack(x, f, ack, true, now);
            return;
        }
        switch (state)
        {
        case C_OPEN:
            cc_acked(x, s, segs, ack, now);
            break;
        case C_DISORDER:
            if (segs >= dups)
                cc_acked(x, s, segs - dups, ack, now);
            break;
        case C_RECOVERY:
            segs = (ack - una_before) / MSS; // (the segments of the whole recovery)
            cc_acked(x, s, segs, ack, now);
            s.first_partial = 1;
            s.dupacks = 0;
            break;
        default: // C_LOSS
            segs = (ack - s.recover) / MSS;
            cc_acked(x, s, segs, ack, now);
            s.first_partial = 1;
            break;
        }
        s.cong = C_OPEN;
        if (ack >= s.recover)
            s.recover_active = 0;
        if (state == C_RECOVERY)
        {
            rto_after_ack(x, f, ack, true, now);
            s.cwnd = s.ssthresh;
        }
        cc_grow(x, s, segs, ack, now);
        rto_after_ack(x, f, ack, true, now);
    }
    // an ACK at the sender (ReceivedAck + ProcessAck, no SACK option, classic recovery): old ones
    // are ignored; the RTT sample and the ledger come first, then the CWR end and the echo, then
    // by kind - a duplicate, an ACK of nothing new at s.high, one beyond it, or new data - and at
    // last the FIN's acknowledgment, the completed messages and more sending
    static __device__ void ack_in(Host &x, int f, uint32_t ack, bool ece, int64_t now)
    {
        SndState &s = *x.ls;
        if (s.st != S_EST || ack < s.una)
            return;
        const uint32_t una_before = s.una;
        const bool advancing = ack > una_before;
        rtt_on_ack(x, f, ack, now);
        Ledger L(*x.B, s, conn(x, f));
        const uint32_t acked = advancing ? lg_release(L, ack) : 0;
        if (s.cong == C_CWR && ack > s.recover)
        {
            s.cong = C_OPEN;
            s.cwnd = s.ssthresh;
        }
        ack_echo(x, s, ack, ece, advancing);
        const uint32_t dups = s.dupacks;
        if (!advancing)
        {
            if (ack < s.high)
            {
                ack_duplicate(x, f, s, una_before, now);
                cc_acked(x, s, 1, ack, now);
            }
            else if (ack > s.high)
                s.nxt = ack;
        }
        else
        {
            // whole segments acknowledged, the remainders carried over
            uint32_t segs = acked / MSS;
            s.ack_rem += acked % MSS;
            if (s.ack_rem >= MSS)
            {
                s.ack_rem -= MSS;
                segs++;
            }
            s.dupacks = 0;
            if (ack < s.recover && (s.cong == C_RECOVERY || s.cong == C_LOSS))
                ack_partial(x, f, s, ack, segs, now);
            else
                ack_advance(x, f, s, ack, una_before, segs, dups, now);
        }
        if ((s.fin & FIN_SENT) && !s.it_n && ack == s.high + 1)
            s.fin |= FIN_ACKED;
        if (advancing)
        {
            msgs_complete(x, f, now);
            if (s.st == S_DONE)
                return;
        }
        tx_fill(x, f, now);
    }
    // the retransmission timeout: every outstanding segment lost, the timeout doubled (to 60 s at
    // most), one segment's window, then sending again
    static __device__ void rto_fire(Host &x, int f, int64_t now)
    {
        SndState &s = *x.ls;
        s.rto_at = -1;
        if (s.st != S_EST)
            return;
        if (s.una >= s.tail)
        {
            if ((s.fin & (FIN_SENT | FIN_ACKED)) == FIN_SENT)
                count(x.R, EV_RTT_OVF);
            return;
        }
        count(x.R, EV_RTO);
        const uint32_t was_in_flight = inflight(s);
        Ledger L(*x.B, s, conn(x, f));
        lg_all_lost(L);
        s.dupacks = 0;
        s.recover = s.high;
        s.recover_active = 1;
        s.rto = 2 * s.rto < 60000000000000ll ? 2 * s.rto : 60000000000000ll;
        s.h_n = 0;
        s.ssthresh = cc_ssthresh(x, s, was_in_flight);
        if (x.cubic)
            cubic_timeout(s, now);
        s.cong = C_LOSS;
        s.cwnd = MSS;
        tx_fill(x, f, now);
    }

    // -------------------------------------------------------------- the receiving endpoint
    // A connection's receiving side: the SYN-ACK and its retries, in-order delivery through the
    // reassembly ring, delayed ACKs (every second segment or after the delay), DCTCP's CE echo
    // state machine, and the passive close (FIN+ACK, LAST_ACK with its retries).
    // an ACK now (the delayed one, if pending, is folded in); ECE while echoing CE unless no_ece,
    // always with force_ece
    static __device__ void rcv_ack(Host &x, int f, uint32_t ack, bool force_ece, bool no_ece, int64_t now)
    {
        RcvState &r = *x.lr;
        const bool echoing = r.ecn_r == E_CE_RCVD || r.ecn_r == E_SENDING_ECE;
        const bool ece = force_ece || (!no_ece && echoing);
        if (!force_ece && ece)
            r.ecn_r = E_SENDING_ECE;
        r.delack_at = -1;
        r.delack_count = 0;
        emit(x, f, (uint8_t)(F_ACK | (ece ? F_ECE : 0)), 0, ack, CTRL_FRAME, true, now);
    }
    static __device__ void rcv_synack(Host &x, int f, int64_t now)
    {
        RcvState &r = *x.lr;
        const int64_t backoff = CONN_TIMEOUT_PS << r.srv_syn_n++;
        if (r.r_synack_t < 0)
            r.r_synack_t = now;
        emit(x, f, F_SYN | F_ACK, 0, 0, SYN_FRAME, true, now);
        if (r.srv_syn_at >= 0)
            return;
        r.srv_syn_at = now + backoff;
        timer_chain(x, x.B->dchain, f, backoff);
    }
    static __device__ __forceinline__ void rcv_established(RcvState &r)
    {
        if (r.r_est)
            return;
        r.r_est = 1;
        r.srv_syn_at = -1;
    }
    // DCTCP's receiver: a change of the CE state flushes a pending delayed ACK for the bytes
    // before it, with the previous state's ECE (set when leaving CE, cleared when entering it)
    static __device__ void rcv_ce_edge(Host &x, int f, bool ce, int64_t now)
    {
        RcvState &r = *x.lr;
        const bool flips = (bool)r.ce_state != ce;
        if (flips && r.reserved && r.prior_flag)
            rcv_ack(x, f, r.prior_rcv_nxt, !ce, ce, now);
        r.ce_state = ce ? 1 : 0;
        if (ce)
            r.ecn_r = E_CE_RCVD;
        else if (r.ecn_r == E_CE_RCVD || r.ecn_r == E_SENDING_ECE)
            r.ecn_r = E_IDLE;
        r.prior_flag = 1;
        r.prior_rcv_nxt = r.rcv_nxt;
    }
    static __device__ void rcv_ecn_in(Host &x, int f, const Frame &p, int64_t now)
    {
        if (!x.dctcp)
            return;
        RcvState &r = *x.lr;
        if ((p.flags & F_CE) && r.ecn_ce_seq < (int64_t)p.sa)
        {
            r.ecn_ce_seq = p.sa;
            r.ecn_r = E_CE_RCVD;
            rcv_ce_edge(x, f, true, now);
        }
        else if (p.flags & F_ECT)
            rcv_ce_edge(x, f, false, now);
    }
    // the CE state of the sending side's own ECN (ACKs carrying CE)
    static __device__ void snd_ecn_in(Host &x, SndState &s, const Frame &p)
    {
        if (!x.dctcp)
            return;
        if ((p.flags & F_CE) && !(s.s_ce & 1))
        {
            s.s_ce = 3;
            s.ecn_s = E_CE_RCVD;
            return;
        }
        if (!(p.flags & F_ECT))
            return;
        s.s_ce &= ~2;
        if (s.ecn_s == E_CE_RCVD || s.ecn_s == E_SENDING_ECE)
            s.ecn_s = E_IDLE;
    }
    // The reassembly ring (TcpRxBuffer::Add): out-of-order blocks, disjoint and sorted, in a ring
    // per connection. [seq, end) above rcv_nxt joins the blocks it touches; a block reaching
    // rcv_nxt is delivered. False if nothing was new (or the ring is full).
    static __device__ bool reasm_add(Host &x, int f, uint32_t seq, uint32_t end)
    {
        const Batch &B = *x.B;
        RcvState &r = *x.lr;
        const ConnStatic &c = conn(x, f);
        const uint32_t from = seq > r.rcv_nxt ? seq : r.rcv_nxt;
        if (from >= end)
            return false;
        uint32_t *lo = B.oo_s + c.oo_off, *hi = B.oo_e + c.oo_off;
        const uint32_t cap = c.oo_cap, h0 = r.ooo_h;
        uint32_t n = r.ooo_n;
        auto at = [&](uint32_t i) { return h0 + i < cap ? h0 + i : h0 + i - cap; };
        // the first block ending at or after `from`
        uint32_t i = n;
        if (n && hi[at(n - 1)] >= from)
        {
            uint32_t a = 0, b = n - 1;
            while (a < b)
            {
                const uint32_t m = (a + b) / 2;
                if (hi[at(m)] < from)
                    a = m + 1;
                else
                    b = m;
            }
            i = a;
        }
        if (i < n && lo[at(i)] <= from && hi[at(i)] >= end)
            return false; // (all of it is held already)
        // blocks i .. j - 1 touch [from, end): they merge into one
        uint32_t j = i, nlo = from, nhi = end;
        for (; j < n && lo[at(j)] <= nhi; j++)
        {
            nlo = lo[at(j)] < nlo ? lo[at(j)] : nlo;
            nhi = hi[at(j)] > nhi ? hi[at(j)] : nhi;
        }
        if (j == i)
        {
            // a new block at i
            if (n >= cap)
            {
                count(x.R, EV_OOO_OVF);
                return false;
            }
            for (uint32_t k = n; k > i; k--)
            {
                lo[at(k)] = lo[at(k - 1)];
                hi[at(k)] = hi[at(k - 1)];
            }
            n++;
        }
        else if (j - i > 1)
        {
            // the merged blocks after the first close up
            const uint32_t gone = j - i - 1;
            for (uint32_t k = j; k < n; k++)
            {
                lo[at(k - gone)] = lo[at(k)];
                hi[at(k - gone)] = hi[at(k)];
            }
            n -= gone;
        }
        lo[at(i)] = nlo;
        hi[at(i)] = nhi;
        uint32_t head = h0;
        if (from == r.rcv_nxt)
        {
            // the first block is in order now: delivered
            r.rcv_nxt = hi[at(0)];
            head = at(1);
            n--;
        }
        r.ooo_n = n;
        r.ooo_h = n ? head : 0;
        if (r.r_fin >= 0 && r.rcv_nxt == (uint32_t)r.r_fin)
            r.rcv_nxt++; // (the FIN's sequence number)
        return true;
    }
    static __device__ __forceinline__ bool rcv_fin_in(const RcvState &r) { return r.r_fin >= 0 && (uint32_t)r.r_fin < r.rcv_nxt; }
    static __device__ __forceinline__ int64_t rcv_rto(const RcvState &r)
    {
        const int64_t base = r.r_srtt >= 0 ? r.r_srtt : INIT_RTT_PS, spread = 4 * r.r_rttvar > CLOCK_GRAN_PS ? 4 * r.r_rttvar : CLOCK_GRAN_PS;
        return base + spread;
    }
    // FIN+ACK of the passive close, with its retransmission timer (at least MinRto) if none runs
    static __device__ void rcv_fin_ack(Host &x, int f, int64_t now)
    {
        RcvState &r = *x.lr;
        r.delack_at = -1;
        r.delack_count = 0;
        emit(x, f, F_ACK | F_FIN, 0, r.rcv_nxt, CTRL_FRAME, true, now);
        if (r.r_frx_at >= 0)
            return;
        const int64_t t = rcv_rto(r) > MIN_RTO_PS ? rcv_rto(r) : MIN_RTO_PS;
        r.r_frx_at = now + t;
        timer_chain(x, x.B->fchain, f, t);
    }
    // the peer closed: LAST_ACK, our FIN+ACK, and the LastAck timer with its retries
    static __device__ void rcv_peer_closed(Host &x, int f, int64_t now)
    {
        RcvState &r = *x.lr;
        r.r_st = R_LAST_ACK;
        rcv_fin_ack(x, f, now);
        r.r_last_n = DATA_RETRIES;
        const int64_t t = rcv_rto(r);
        r.r_last_at = now + t;
        timer_chain(x, x.B->dchain, f, t);
    }
    static __device__ __forceinline__ void rcv_closed(RcvState &r)
    {
        r.r_st = R_CLOSED;
        r.delack_at = r.r_last_at = r.r_frx_at = -1;
    }
    // data delivered to the ring: an immediate ACK for anything out of order (a gap left, or one
    // filled), for nothing new, or for every second segment; else the delayed ACK
    static __device__ void rcv_data(Host &x, int f, uint32_t seq, uint32_t len, bool fin, int64_t now)
    {
        RcvState &r = *x.lr;
        const uint32_t before = r.rcv_nxt;
        if (!reasm_add(x, f, seq, seq + len))
        {
            rcv_ack(x, f, r.rcv_nxt, false, false, now);
            return;
        }
        if (!fin && before < r.rcv_nxt && rcv_fin_in(r))
        {
            rcv_peer_closed(x, f, now);
            return;
        }
        const bool irregular = r.ooo_n > 0 || r.rcv_nxt > before + len;
        const uint32_t held = r.delack_count + 1u;
        if (irregular || held >= DELACK_COUNT)
        {
            r.reserved = 0;
            rcv_ack(x, f, r.rcv_nxt, false, false, now);
            return;
        }
        r.delack_count = (uint8_t)held;
        r.reserved = 1;
        if (r.delack_at >= 0)
            return;
        r.delack_at = now + DELACK_PS;
        timer_chain(x, x.B->dchain, f, DELACK_PS);
    }
    // a data (or FIN) segment at the receiver: old data in an established connection is answered
    // with an ACK (in LAST_ACK: anything but the next byte); the CE state; a CWR ends the echo;
    // the first segment establishes; in LAST_ACK only the close's ACK counts; a FIN is taken in
    // order, marks the end and closes passively once delivered
    static __device__ void rcv_segment(Host &x, int f, const Frame &p, int64_t now)
    {
        RcvState &r = *x.lr;
        if (!r.r_up || r.r_st == R_CLOSED)
            return;
        const bool fin = p.flags & F_FIN;
        const uint32_t seq = p.sa, len = (p.flags & F_DATA) ? p.len - 42u : 0u, end = seq + len;
        if (len && r.r_est)
        {
            const bool stale = r.r_st == R_LAST_ACK ? seq != r.rcv_nxt : end < r.rcv_nxt;
            if (stale)
            {
                rcv_ack(x, f, r.rcv_nxt, false, true, now);
                return;
            }
        }
        rcv_ecn_in(x, f, p, now);
        if ((p.flags & F_CWR) && r.r_est && r.r_st == R_OPEN && r.ecn_r != E_CE_RCVD)
            r.ecn_r = E_IDLE;
        if (!r.r_est)
        {
            if (fin && seq != r.rcv_nxt)
                return;
            rcv_established(r);
        }
        if (r.r_st == R_LAST_ACK)
        {
            if (fin || seq == r.rcv_nxt)
                rcv_closed(r);
            return;
        }
        if (!fin)
        {
            rcv_data(x, f, seq, len, false, now);
            return;
        }
        if (seq < r.rcv_nxt)
            return;
        r.r_fin = end;
        if (r.rcv_nxt == end)
            r.rcv_nxt++;
        if (len)
            rcv_data(x, f, seq, len, true, now);
        if (rcv_fin_in(r))
            rcv_peer_closed(x, f, now);
    }

    // -------------------------------------------------------------- application and dispatch
    static __device__ void add_active(Host &x, bool snd, int f)
    {
        int32_t *n = snd ? x.B->n_snd : x.B->n_rcv;
        const int k = n[x.h];
        if (k >= MAX_ACTIVE)
        {
            count(x.R, EV_ACTIVE_OVF);
            return;
        }
        (snd ? x.B->act_snd : x.B->act_rcv)[(size_t)x.h * MAX_ACTIVE + k] = f;
        n[x.h] = k + 1;
    }
    // the application writes its due messages (FlowSender::Write); Socket::Send defers
    // SendPendingData by 1 ps when the window has room (src/step.cuh:2120-2145)
    static __device__ void app_write(Host &x, int f, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        const ConnStatic &c = conn(x, f);
        if (s.m_wr >= s.m_due)
            return;
        for (; s.m_wr < s.m_due; s.m_wr++)
        {
            const int m = c.m_first + s.m_wr;
            B.mg_est[m] = now;
            s.tail = B.mg_end[m];
        }
        if (c.fin_on && s.m_wr == c.m_cnt)
            s.fin |= FIN_CLOSE;
        if (s.cwnd <= inflight(s) || s.syn_at >= 0)
            return;
        s.syn_at = now + 1;
        Chain ch;
        chain_child(ch, 2, x.echain);
        chain_store(B.schain + (size_t)f * CW, ch);
        due_insert(x, now + 1, f, 2);
    }
    // a frame at its destination host (src/step.cuh:2147-2236): the sending side's part - the
    // SYN-ACK establishes (RTT sample, the handshake's ACK, the first write), a FIN+ACK (the
    // sender's close) is acknowledged once, any other ACK goes to ack_in
    static __device__ void on_frame_snd(Host &x, int f, const Frame &p, int64_t now)
    {
        SndState &s = *x.ls;
        if (p.flags & F_SYN)
        {
            if (s.st != S_SYN_SENT)
                return;
            s.st = S_EST;
            s.syn_at = -1;
            rtt_sample(s, now - s.timing_time);
            emit(x, f, F_ACK, 0, 0, CTRL_FRAME, false, now);
            chain_child(x.echain, 1, x.echain);
            app_write(x, f, now);
            return;
        }
        if (s.fin & FIN_TW)
            return;
        snd_ecn_in(x, s, p);
        ack_in(x, f, p.sa, p.flags & F_ECE, now);
        if (!(p.flags & F_FIN))
            return;
        s.fin |= FIN_TW;
        emit(x, f, F_ACK, 0, 1, CTRL_FRAME, false, now);
    }
    // ... and the receiving side's: the handshake's RTT for the receiver, a SYN (the connection
    // comes up; a SYN-ACK unless established), data, or an ACK (establishes; ends LAST_ACK)
    static __device__ void on_frame_rcv(Host &x, int f, const Frame &p, int64_t now)
    {
        RcvState &r = *x.lr;
        const bool syn = p.flags & F_SYN;
        if (!syn && r.r_up && r.r_srtt < 0 && r.r_synack_t >= 0)
        {
            const int64_t m = now - r.r_synack_t;
            r.r_srtt = m;
            r.r_rttvar = m / 2;
        }
        if (syn)
        {
            if (!r.r_up)
            {
                r.r_up = 1;
                r.delack_count = DELACK_COUNT;
                add_active(x, false, f);
            }
            if (r.r_est)
                return;
            x.nic.woke = now;
            chain_child(x.echain, 1, x.echain);
            rcv_synack(x, f, now);
            return;
        }
        if (p.flags & (F_DATA | F_FIN))
        {
            rcv_segment(x, f, p, now);
            return;
        }
        if (!r.r_up || r.r_st == R_CLOSED)
            return;
        rcv_ecn_in(x, f, p, now);
        rcv_established(r);
        if (r.r_st == R_LAST_ACK)
            rcv_closed(r);
    }
    // one event per frame: the side's record in LDS for its duration
    static __device__ void on_frame(Host &x, const Frame &p, int64_t now)
    {
        const int f = (int)p.flow;
        if (p.to_sender)
        {
            snd_open(x, f);
            on_frame_snd(x, f, p, now);
            snd_close(x, f);
        }
        else
        {
            rcv_open(x, f);
            on_frame_rcv(x, f, p, now);
            rcv_close(x, f);
        }
    }
    static __device__ __forceinline__ const uint32_t *due_chain(const Batch &B, const Due &d)
    {
        const uint32_t *base = d.kind == 1 || d.kind == 3 || d.kind == 4 ? B.dchain
                               : d.kind == 5                            ? B.fchain
                               : (d.kind == 2 && B.snd[d.flow].st == S_EST) ? B.schain
                                                                            : B.rchain;
        return base + (size_t)d.flow * CW;
    }
    // a timer and an arrival at one instant: the one scheduled first runs first (undecided: the timer)
    static __device__ __forceinline__ bool timer_before_arrival(Host &x, const Due &d, const Frame *p)
    {
        Chain arr, pc;
        msg_chain(pc, p);
        chain_child(arr, 2048u + msg_head(p).len, pc);
        int r = chain_cmp(due_chain(*x.B, d), arr.r);
        if (r == 0)
            count(x.R, EV_UNDECIDED);
        return r <= 0;
    }

    // -------------------------------------------------------------- one slot of a host
    // (src/step.cuh:2263-2528): returns 1 if it sent or holds frames; *wake: its next slot
    static __device__ int host_step(const Batch &B, const WorldDesc &W, WorldRun *R, int hl, int64_t t, int64_t *wake, SndState *ls,
                                    RcvState *lr, ConnStatic *lc)
    {
        const int64_t S0 = t * SLOT_PS, E0 = S0 + SLOT_PS;
        const int h = W.host0 + hl, lid = W.link0 + hl;
#ifdef TW_PROF
        const unsigned long long pr0 = wall_clock64();
        unsigned long long pr_ev = 0, pr_n = 0;
#endif
        const int in = B.in_link[lid];
        const Frame *m2 = nullptr, *m1 = nullptr;
        int n2 = 0, n1 = 0, i2 = 0, i1 = 0;
        // this slot's arrivals: the later part of slot t-2's message, the earlier of t-1's
        if (t >= 2)
        {
            size_t r = (size_t)((t - 2) % NRING) * B.nlinks + in;
            const long long v = coh_load(B.msg_cnt + r);
            if (msg_slot(v) == t - 2)
                i2 = msg_early(v), n2 = msg_sent(v);
            m2 = B.msg + r * MAX_TX;
        }
        if (t >= 1)
        {
            size_t r = (size_t)((t - 1) % NRING) * B.nlinks + in;
            const long long v = coh_load(B.msg_cnt + r);
            n1 = msg_slot(v) == t - 1 ? msg_early(v) : 0;
            m1 = B.msg + r * MAX_TX;
        }
        bool input = i2 < n2 || n1 > 0;
        if (!input && t < B.wake[h] && B.q_size[lid] == 0)
        {
            *wake = B.wake[h];
            return 0;
        }
        const int rt = (int)(t % NRING);
        Host x;
        x.nic.load(B, R, lid, t, rt);
        x.B = &B;
        x.W = &W;
        x.R = R;
        x.h = h;
        x.hl = hl;
        x.dctcp = B.host_cc[h] == CC_DCTCP;
        x.cubic = B.host_cc[h] == CC_CUBIC;
        x.due = B.due + (size_t)h * MAX_DUE;
        x.slot_end = E0;
        x.ls = ls;
        x.lr = lr;
        x.lc = lc;
        Due *due = x.due;
        // timers due in this slot, by time. A list is scanned only when one of its timers falls in
        // the slot: tmin_s / tmin_r are its earliest, exact since the end of the host's last slot
        // (timers change only while the host runs)
        int nd = 0;
        const int64_t bound_s = B.tmin_s[h], bound_r = B.tmin_r[h];
        const bool scan_s = bound_s < E0, scan_r = bound_r < E0;
        const int ns = scan_s ? B.n_snd[h] : 0, nr = scan_r ? B.n_rcv[h] : 0;
        const int32_t *as = B.act_snd + (size_t)h * MAX_ACTIVE, *ar = B.act_rcv + (size_t)h * MAX_ACTIVE;
        for (int i = 0; i < ns; i++)
        {
            int f = as[i];
            int64_t a = B.snd[f].rto_at, b = B.snd[f].syn_at;
            if (a >= 0 && a < E0 && nd < MAX_DUE)
                due[nd++] = {a, f, 0};
            if (b >= 0 && b < E0 && nd < MAX_DUE)
                due[nd++] = {b, f, 2};
        }
        for (int i = 0; i < nr; i++)
        {
            int f = ar[i];
            const RcvState &r = B.rcv[f];
            int64_t a = r.delack_at, b = r.srv_syn_at, l = r.r_last_at, xx = r.r_frx_at;
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
            count(R, EV_DUE_OVF);
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
        x.due_head = nd ? due[0].at : INT64_MAX;
        int pn = B.pend_next[h];
        const int pe = B.pend_end[h];
        // The three candidates' times are kept between events, each read again only when its
        // own event is taken (the next reads are then in flight while that event runs): the
        // next arrival, the next timer (due_head, which due_insert keeps), the next application
        // start of the host's order.
        // (the next arrival's whole frame is read with its time, so taking it waits for no load)
        Frame fa;
        auto next_arrival = [&]() -> int64_t {
            if (i2 >= n2 && i1 >= n1)
                return INT64_MAX;
            fa = coh_get(i2 < n2 ? m2 + i2 : m1 + i1);
            return fa.ts;
        };
        auto next_start = [&](int &m) -> int64_t {
            m = pn < pe ? B.pend_msg[pn] : -1;
            if (m >= 0)
            {
                const int64_t at = B.mg_start[m];
                if (at < E0)
                    return at;
            }
            return INT64_MAX;
        };
        int pm;
        int64_t tr = next_arrival(), tp = next_start(pm);
#ifdef TW_PROF
        const unsigned long long pr1 = wall_clock64();
#endif
        for (;;)
        {
            const int64_t td = x.due_head;
            int64_t now = tr < td ? tr : td;
            now = now < tp ? now : tp;
            if (now == INT64_MAX)
                break;
            if (now < S0)
                now = S0;
            x.nic.advance(now);
#ifdef TW_PROF
            const unsigned long long pq = wall_clock64();
            pr_n++;
#endif
            if (now == tp && B.cs[B.mg_conn[pm]].m_first != pm)
            {
                // a later message of a connection (FlowSender::SendMessage)
                const int m = pm;
                pn++;
                tp = next_start(pm);
                int f = B.mg_conn[m];
                chain_root_msg(x.echain, m - W.msg0);
                SndState &s = snd_open(x, f);
                s.m_due++;
                if (s.st == S_EST)
                    app_write(x, f, now);
                snd_close(x, f);
            }
            else if (now == tp)
            {
                // a connection's application starts: SYN
                const int m = pm;
                pn++;
                tp = next_start(pm);
                int f = B.mg_conn[m];
                SndState &s = snd_open(x, f);
                chain_root(x.echain, hl, f - W.conn0);
                s.m_due = 1;
                add_active(x, true, f);
                s.st = S_SYN_SENT;
                if (x.cubic)
                    cubic_round_start(s, now);
                s.rto = CONN_TIMEOUT_PS;
                s.syn_at = now + CONN_TIMEOUT_PS;
                timer_chain(x, B.rchain, f, CONN_TIMEOUT_PS);
                s.timing_time = now;
                emit(x, f, F_SYN, 0, 0, SYN_FRAME, false, now);
                snd_close(x, f);
            }
            else if (now == td && (now != tr || timer_before_arrival(x, due[x.id], i2 < n2 ? m2 + i2 : m1 + i1)))
            {
                Due d = due[x.id++];
                x.due_head = x.id < x.nd ? due[x.id].at : INT64_MAX;
                chain_load(x.echain, due_chain(B, d));
                const int f = d.flow;
                if (d.kind == 0)
                {
                    if (B.snd[f].rto_at == d.at)
                    {
                        snd_open(x, f);
                        rto_fire(x, f, now);
                        snd_close(x, f);
                    }
                }
                else if (d.kind == 2)
                {
                    SndState &s = snd_open(x, f);
                    if (s.syn_at == d.at && s.st == S_EST)
                    {
                        s.syn_at = -1; // the deferred SendPendingData
                        tx_fill(x, f, now);
                    }
                    else if (s.syn_at == d.at && s.st == S_SYN_SENT)
                    {
                        int tries = ++s.syn_tries; // SYN retransmission, backed off
                        s.rto = CONN_TIMEOUT_PS << tries;
                        s.syn_at = now + s.rto;
                        timer_chain(x, B.rchain, f, s.rto);
                        emit(x, f, F_SYN, 0, 0, SYN_FRAME, false, now);
                    }
                    snd_close(x, f);
                }
                else
                {
                    RcvState &r = rcv_open(x, f);
                    if (d.kind == 1 && r.delack_at == d.at)
                    {
                        r.delack_at = -1;
                        r.delack_count = 0;
                        r.reserved = 1;
                        rcv_ack(x, f, r.rcv_nxt, false, false, now);
                    }
                    else if (d.kind == 4 && r.r_last_at == d.at && r.r_st == R_LAST_ACK)
                    {
                        if (r.r_last_n == 0)
                            rcv_closed(r);
                        else
                        {
                            r.r_last_n--;
                            rcv_fin_ack(x, f, now);
                            int64_t lr = rcv_rto(r);
                            r.r_last_at = now + lr;
                            timer_chain(x, B.dchain, f, lr);
                        }
                    }
                    else if (d.kind == 5 && r.r_frx_at == d.at && r.r_st == R_LAST_ACK)
                    {
                        r.r_frx_at = -1;
                        rcv_fin_ack(x, f, now);
                    }
                    else if (d.kind == 3 && r.srv_syn_at == d.at && !r.r_est)
                    {
                        r.srv_syn_at = -1;
                        rcv_synack(x, f, now);
                    }
                    rcv_close(x, f);
                }
            }
            else
            {
                const Frame p = fa;
                if (i2 < n2)
                    i2++;
                else
                    i1++;
                tr = next_arrival();
                Chain pc;
                chain_load(pc, p.chain);
                chain_child(x.echain, 2048u + p.len, pc);
                on_frame(x, p, now);
            }
#ifdef TW_PROF
            pr_ev += wall_clock64() - pq;
#endif
        }
#ifdef TW_PROF
        const unsigned long long pr2 = wall_clock64();
#endif
        x.nic.advance(E0 - 1);
        x.nic.store(t, rt);
        B.pend_next[h] = pn;
        // drop finished connections from the active lists; their earliest timers; the next wake-up
        const int64_t nb_s = compact_snd(B, h), nb_r = compact_rcv(B, h);
        B.tmin_s[h] = nb_s;
        B.tmin_r[h] = nb_r;
        int64_t next = nb_s < nb_r ? nb_s : nb_r;
        if (pn < pe && B.mg_start[B.pend_msg[pn]] < next)
            next = B.mg_start[B.pend_msg[pn]];
        int64_t wk = x.nic.size > 0 ? t + 1 : (next == INT64_MAX ? INT64_MAX : next / SLOT_PS);
        B.wake[h] = wk;
        *wake = wk;
#ifdef TW_PROF
        const unsigned long long pr3 = wall_clock64();
        atomicAdd(&R->ev[EV_PROF_H_SETUP], pr1 - pr0);
        atomicAdd(&R->ev[EV_PROF_H_LOOP], pr2 - pr1);
        atomicAdd(&R->ev[EV_PROF_H_EVENTS], pr_ev);
        atomicAdd(&R->ev[EV_PROF_H_END], pr3 - pr2);
        atomicAdd(&R->ev[EV_PROF_H_COUNT], pr_n);
#endif
        return (x.nic.sent > 0 || x.nic.size > 0) ? 1 : 0;
    }

    // -------------------------------------------------------------- routing (src/step.cuh:2541-2630)
    static __device__ int epoch_of(const Batch &B, const WorldDesc &W, int64_t ts)
    {
        int e = 0;
        while (e < W.n_fail && ts >= B.g_fail_t[W.o_fail + e])
            e++;
        return e;
    }
    static __device__ __forceinline__ size_t rt_index(const WorldDesc &W, int ep, int x, int e)
    {
        return (size_t)W.o_rt + ((size_t)ep * W.S + x) * W.S + e;
    }
    // the output port at switch sw (world-local) towards host dst; up: the equal-cost index
    static __device__ int route_up(const Batch &B, const WorldDesc &W, int sw, int dst, int up, int ep)
    {
        if (W.kind == TOPO_GRAPH)
        {
            int e = B.g_host_sw[W.o_host + dst];
            if (e == sw)
                return B.g_host_port[W.o_host + dst];
            return B.g_rt_ports[W.o_rtp + B.g_rt_off[rt_index(W, ep, sw, e)] + up];
        }
        if (W.kind == TOPO_LEAFSPINE)
        {
            if (sw < W.leaves)
                return dst / W.hpl == sw ? dst % W.hpl : W.hpl + up;
            return dst / W.hpl;
        }
        const int hk = W.hk, E = W.k * hk;
        int dst_pod = dst / (hk * hk), dst_edge = (dst / hk) % hk;
        if (sw < E)
            return (dst_pod == sw / hk && dst_edge == sw % hk) ? dst % hk : hk + up;
        if (sw < 2 * E)
            return dst_pod == (sw - E) / hk ? dst_edge : hk + up;
        return dst_pod;
    }
    static __device__ int route(const Batch &B, const WorldDesc &W, int sw, const FrameHead &p, int ep)
    {
        if (W.kind == TOPO_GRAPH)
        {
            int e = B.g_host_sw[W.o_host + p.dst];
            if (e == sw)
                return B.g_host_port[W.o_host + p.dst];
            const ConnStatic &c = B.cs[p.flow];
            uint32_t hh = p.to_sender ? c.hash_ack : c.hash_data;
            return route_up(B, W, sw, p.dst, (int)(hh % B.g_rt_n[rt_index(W, ep, sw, e)]), ep);
        }
        return route_up(B, W, sw, p.dst, p.up, ep);
    }
    static __device__ __forceinline__ bool port_down(const Batch &B, const WorldDesc &W, int sw, int q, int ep)
    {
        return W.n_fail && B.g_port_dn[W.o_pdn + ((size_t)ep * W.S + sw) * W.k + q];
    }
    static __device__ int routes_to(const Batch &B, const WorldDesc &W, int sw, int dst, int ep)
    {
        if (W.kind == TOPO_GRAPH)
        {
            int e = B.g_host_sw[W.o_host + dst];
            return e == sw ? 0 : B.g_rt_n[rt_index(W, ep, sw, e)];
        }
        if (W.kind == TOPO_LEAFSPINE)
            return sw < W.leaves ? (dst / W.hpl == sw ? 0 : W.ecmp_n) : 1;
        const int hk = W.hk, E = W.k * hk;
        int dst_pod = dst / (hk * hk), dst_edge = (dst / hk) % hk;
        if (sw < E)
            return dst_pod == sw / hk && dst_edge == sw % hk ? 0 : W.ecmp_n;
        if (sw < 2 * E)
            return dst_pod == (sw - E) / hk ? 1 : W.ecmp_n;
        return 1;
    }
    // simultaneous arrivals: the longer frame first, then the upstream starts' chains, then
    // (undecided) the flow id (src/step.cuh:2691-2704)
    static __device__ bool frame_before(WorldRun *R, const Frame *a, const Frame *b)
    {
        const FrameHead ha = msg_head(a), hb = msg_head(b);
        if (ha.len != hb.len)
            return ha.len > hb.len;
        int c = chain_cmp_by([&](int i) { return coh_load(&a->chain[i]); }, [&](int i) { return coh_load(&b->chain[i]); });
        if (c == 0)
            count(R, EV_UNDECIDED);
        return c != 0 ? c < 0 : ha.flow < hb.flow;
    }

    // -------------------------------------------------------------- one slot of a switch
    // (src/step.cuh:2706-2886). The k lanes of one switch within one wavefront, lane q = port q;
    // phases separated by wave_sync(). LDS per switch: rows[k][k+1] counts, in_win[k][3], base[k],
    // cnt[k]; per-wave global scratch holds the arrivals sorted per egress.
    // a switch item's LDS: per lane rows (33 ints), in_win (3), base, cnt, then (after the
    // spray merge's 64) per lane SORT_LDS arrival times and refs for the egress sort, then per
    // lane the egress ports of its first EGRESS_LDS arrivals (one byte each: the count phase
    // routes them, the scatter phase reads the routes back), then per lane its port's in-link
    static constexpr int SW_LDS_INTS = 64 * 33 + 64 * 3 + 64 + 64, SORT_LDS = 24, EGRESS_LDS = 64;
    static constexpr int SW_LDS_EGRESS = SW_LDS_INTS + 64 * SORT_LDS * 3;
    static constexpr int SW_LDS_INLINK = SW_LDS_EGRESS + 64 * EGRESS_LDS / 4;
    static constexpr int SW_LDS_TOTAL = SW_LDS_INLINK + 64;
    static constexpr uint8_t NO_EGRESS = 0xFF;
    struct SwLds
    {
        int *rows, *in_win, *base, *cnt;
        int *wave; // the wavefront's whole LDS region (33 ints per lane at the start)
    };
    struct In
    {
        const Frame *m2, *m1;
        int i2, n2, n1;
        __device__ int count() const { return (n2 - i2) + n1; }
        // the j-th arrival (a frame another entity wrote in an earlier slot: read it device-coherently)
        __device__ const Frame *at(int j) const { return j < n2 - i2 ? m2 + i2 + j : m1 + (j - (n2 - i2)); }
        __device__ FrameHead head(int j) const { return msg_head(at(j)); }
    };
    static __device__ In input_of(const Batch &B, int link, int64_t t, const int *win)
    {
        In in;
        in.i2 = win[0];
        in.n2 = win[1];
        in.n1 = win[2];
        if (link < 0)
            link = 0; // (an unused port has an empty window)
        in.m2 = B.msg + ((size_t)((t + NRING - 2) % NRING) * B.nlinks + link) * MAX_TX;
        in.m1 = B.msg + ((size_t)((t + NRING - 1) % NRING) * B.nlinks + link) * MAX_TX;
        return in;
    }
    // One lane of a switch's slot: port q of world-local switch swl (inactive lanes idle).
    // The phases run in order for all lanes of a switch, with wave_sync() between them (a CPU
    // check may run each phase over the lanes one after another). Scratch: this switch's region
    // (k * 2 * MAX_TX entries) of the wave's arrival scratch.
    struct SwLane
    {
        int swl, q, lid;
        bool active;
        int64_t *scr_ts;
        uint32_t *scr_ref;
        uint8_t *scr_up;
    };
    // port q's input of the slot (its in-link from LDS, where the window phase put it)
    static __device__ In lane_input(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t t, const SwLds &L, int q)
    {
        return input_of(B, L.wave[SW_LDS_INLINK + sw_lane_of(L, x) - x.q + q], t, L.in_win + 3 * q);
    }
    // the switch lane's place in its wave (its switch's first lane, where L.cnt points, plus its port)
    static __device__ __forceinline__ int sw_lane_of(const SwLds &L, const SwLane &x)
    {
        return (int)(L.cnt - L.wave - (64 * 33 + 64 * 3 + 64)) + x.q;
    }
    static __device__ __forceinline__ uint8_t *lane_egress_lds(const SwLds &L, const SwLane &x)
    {
        return reinterpret_cast<uint8_t *>(L.wave + SW_LDS_EGRESS) + sw_lane_of(L, x) * EGRESS_LDS;
    }
    static __device__ int lane_egress(const Batch &B, const WorldDesc &W, const SwLane &x, int j, const FrameHead &p)
    {
        int ep = W.n_fail ? epoch_of(B, W, p.ts) : 0;
        if (port_down(B, W, x.swl, x.q, ep))
            return -1; // dropped on arrival at a down interface
        return W.rng0 >= 0 ? route_up(B, W, x.swl, p.dst, x.scr_up[x.q * 2 * MAX_TX + j], ep) : route(B, W, x.swl, p, ep);
    }
    // 1: the ingress window of port q (tail of slot t-2's message from S0 on, head of t-1's
    // before E0) into LDS; the port's row of egress counts cleared
    static __device__ void sw_window(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t t, const SwLds &L)
    {
        if (!x.active)
            return;
        const int in_link = B.in_link[x.lid];
        L.wave[SW_LDS_INLINK + sw_lane_of(L, x)] = in_link;
        int w0 = 0, w1 = 0, w2 = 0;
        if (in_link >= 0)
        {
            if (t >= 2)
            {
                const long long v = coh_load(B.msg_cnt + (size_t)((t - 2) % NRING) * B.nlinks + in_link);
                if (msg_slot(v) == t - 2)
                    w0 = msg_early(v), w1 = msg_sent(v);
            }
            if (t >= 1)
            {
                const long long v = coh_load(B.msg_cnt + (size_t)((t - 1) % NRING) * B.nlinks + in_link);
                w2 = msg_slot(v) == t - 1 ? msg_early(v) : 0;
            }
        }
        L.in_win[3 * x.q] = w0;
        L.in_win[3 * x.q + 1] = w1;
        L.in_win[3 * x.q + 2] = w2;
        int *row = L.rows + x.q * (W.k + 1);
        for (int i = 0; i <= W.k; i++)
            row[i] = 0;
    }
    // 2: arrivals counted by egress (after the spray draws, if any)
    static __device__ void sw_count(const Batch &B, const WorldDesc &W, WorldRun *R, const SwLane &x, int64_t t, const SwLds &L)
    {
        if (!x.active)
            return;
        In in = lane_input(B, W, x, t, L, x.q);
        int *row = L.rows + x.q * (W.k + 1);
        uint8_t *eg = lane_egress_lds(L, x);
        const int n = in.count();
        for (int j = 0; j < n; j++)
        {
            int e = lane_egress(B, W, x, j, in.head(j));
            if (j < EGRESS_LDS)
                eg[j] = e >= 0 ? (uint8_t)e : NO_EGRESS;
            if (e >= 0)
                row[e]++;
            else
                count(R, EV_DROP);
        }
    }
    // 3: for egress q, each ingress's offset in its region and the region's size
    static __device__ void sw_prefix(const WorldDesc &W, const SwLane &x, const SwLds &L)
    {
        if (!x.active)
            return;
        int off = 0;
        for (int j = 0; j < W.k; j++)
        {
            int c = L.rows[j * (W.k + 1) + x.q];
            L.rows[j * (W.k + 1) + x.q] = off;
            off += c;
        }
        L.cnt[x.q] = off;
    }
    // 4: the start of egress q's region: the exclusive prefix sum of the switch's region sizes over
    // its ports. On the GPU the switch's k lanes are consecutive in the wavefront (port q on its
    // first lane + q) and every lane of the wave runs this phase, so a wave64 scan does it in six
    // shuffle steps; a CPU check, which runs the lanes one after another, sums in a loop.
    static __device__ void sw_base(const SwLane &x, const SwLds &L)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const int v = x.active ? L.cnt[x.q] : 0;
        int incl = v;
        for (int d = 1; d < 64; d <<= 1)
        {
            const int u = __shfl_up(incl, d);
            if (x.active && x.q >= d)
                incl += u;
        }
        if (x.active)
            L.base[x.q] = incl - v;
#else
        if (!x.active)
            return;
        int b = 0;
        for (int j = 0; j < x.q; j++)
            b += L.cnt[j];
        L.base[x.q] = b;
#endif
    }
    // 5: every arrival of ingress q into its egress's region, in arrival order
    static __device__ void sw_scatter(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t t, const SwLds &L)
    {
        if (!x.active)
            return;
        In in = lane_input(B, W, x, t, L, x.q);
        int *row = L.rows + x.q * (W.k + 1);
        const uint8_t *eg = lane_egress_lds(L, x);
        const int n = in.count();
        auto scatter = [&](auto *ts, auto *ref) {
            for (int j = 0; j < n; j++)
            {
                const FrameHead p = in.head(j);
                const int e = j < EGRESS_LDS ? (eg[j] == NO_EGRESS ? -1 : (int)eg[j]) : lane_egress(B, W, x, j, p);
                if (e < 0)
                    continue;
                int at = L.base[e] + row[e]++;
                ts[at] = p.ts;
                ref[at] = (uint32_t)(x.q << 16 | j);
            }
        };
        if (sw_lists_in_lds(W, L))
            scatter(sort_ts_lds(L, x), sort_ref_lds(L, x));
        else
            scatter(x.scr_ts, x.scr_ref);
    }
    // the switch's egress lists go to LDS (its lanes' SORT_LDS slots each, laid out by L.base)
    // when all of them fit there, else to the wave's device scratch
    static __device__ __forceinline__ bool sw_lists_in_lds(const WorldDesc &W, const SwLds &L)
    {
        return L.base[W.k - 1] + L.cnt[W.k - 1] <= W.k * SORT_LDS;
    }
    // the switch's sort slots in LDS (from its first lane's)
    static __device__ __forceinline__ int64_t *sort_ts_lds(const SwLds &L, const SwLane &x)
    {
        return reinterpret_cast<int64_t *>(L.wave + SW_LDS_INTS) + (sw_lane_of(L, x) - x.q) * SORT_LDS;
    }
    static __device__ __forceinline__ uint32_t *sort_ref_lds(const SwLds &L, const SwLane &x)
    {
        return reinterpret_cast<uint32_t *>(L.wave + SW_LDS_INTS + 2 * 64 * SORT_LDS) + (sw_lane_of(L, x) - x.q) * SORT_LDS;
    }
    // 6: egress q's arrivals sorted by (time, frame_before), enqueued with the arrival event's
    // chain (RED), and the slot's transmissions; returns 1 if the port sent or holds frames
    static __device__ int sw_egress(const Batch &B, const WorldDesc &W, WorldRun *R, const SwLane &x, int64_t t, const SwLds &L)
    {
        if (!x.active)
            return 0;
        const int n = L.cnt[x.q], b0 = L.base[x.q];
        auto frame_at = [&](uint32_t r) -> const Frame * { return lane_input(B, W, x, t, L, (int)(r >> 16)).at((int)(r & 0xFFFF)); };
        Fifo eg;
        const int rt = (int)(t % NRING);
        eg.load(B, R, x.lid, t, rt);
        // an insertion sort by (time, frame_before): chain comparison is no strict weak order
        // when chains are truncated, and this gives the reference's order in every case; then the
        // arrivals in that order into the queue. (Inlined at two calls, for the list in LDS and for
        // the list in device scratch, so that each copy's accesses have one address space: LDS or
        // global instructions, not flat ones that wait on both.)
        auto sort_enqueue = [&](auto *ts, auto *ref) {
            for (int i = 1; i < n; i++)
            {
                const int64_t xv = ts[i];
                const uint32_t r = ref[i];
                int j = i - 1;
                while (j >= 0 && (ts[j] > xv || (ts[j] == xv && frame_before(R, frame_at(r), frame_at(ref[j])))))
                {
                    ts[j + 1] = ts[j];
                    ref[j + 1] = ref[j];
                    j--;
                }
                ts[j + 1] = xv;
                ref[j + 1] = r;
            }
            for (int i = 0; i < n; i++)
            {
                Frame p = coh_get(frame_at(ref[i]));
                Chain up, c;
                chain_load(up, p.chain);
                chain_child(c, 2048u + p.len, up);
                chain_store(p.chain, c);
                eg.enqueue(p, p.ts);
            }
        };
        if (sw_lists_in_lds(W, L))
            sort_enqueue(sort_ts_lds(L, x) + b0, sort_ref_lds(L, x) + b0); // (the usual case: scattered into LDS)
        else if (n <= SORT_LDS)
        {
            // a short list of a switch whose lists overflow LDS is sorted in the lane's own slots:
            // one round of loads instead of a device-memory round trip per comparison and move
            const int lane = sw_lane_of(L, x);
            int64_t *lts = reinterpret_cast<int64_t *>(L.wave + SW_LDS_INTS) + lane * SORT_LDS;
            uint32_t *lref = reinterpret_cast<uint32_t *>(L.wave + SW_LDS_INTS + 2 * 64 * SORT_LDS) + lane * SORT_LDS;
            for (int i = 0; i < n; i++)
            {
                lts[i] = x.scr_ts[b0 + i];
                lref[i] = x.scr_ref[b0 + i];
            }
            sort_enqueue(lts, lref);
        }
        else
            sort_enqueue(x.scr_ts + b0, x.scr_ref + b0);
        eg.advance((t + 1) * SLOT_PS - 1);
        eg.store(t, rt);
        return (eg.sent > 0 || eg.size > 0) ? 1 : 0;
    }
    // one slot of a switch on the GPU: the lanes of one wavefront, phases separated by wave_sync()
    static __device__ int switch_step(const Batch &B, const WorldDesc &W, WorldRun *R, const SwLane &x, int64_t t, const SwLds &L)
    {
#ifdef TW_PROF
        // profiler builds: the phases' wall-clock ticks, summed over switch items (counters 20-23)
        unsigned long long p0 = wall_clock64();
#define TW_PHASE(i)                                                                                                         \
    {                                                                                                                       \
        unsigned long long p1 = wall_clock64();                                                                             \
        if ((threadIdx.x & 63) == 0)                                                                                        \
            atomicAdd(&R->ev[20 + (i)], p1 - p0);                                                                           \
        p0 = p1;                                                                                                            \
    }
#else
#define TW_PHASE(i)
#endif
        sw_window(B, W, x, t, L);
        if (W.rng0 >= 0)
        {
            wave_sync();
            Spray sp;
            spray_init(B, W, x, t, L, sp);
            wave_sync();
            for (;;)
            {
                spray_put(x, L, sp);
                wave_sync();
                const bool done = spray_pick(B, W, R, x, t, L, sp);
                wave_sync();
                if (done)
                    break;
            }
            spray_end(B, W, x, L, sp);
        }
        wave_sync();
        TW_PHASE(0);
        sw_count(B, W, R, x, t, L);
        wave_sync();
        sw_prefix(W, x, L);
        wave_sync();
        sw_base(x, L);
        wave_sync();
        TW_PHASE(1);
        sw_scatter(B, W, x, t, L);
        wave_sync();
        TW_PHASE(2);
        const int r = sw_egress(B, W, R, x, t, L);
        TW_PHASE(3);
#undef TW_PHASE
        return r;
    }
    // Per-packet ECMP (src/step.cuh:2706-2745): one GetInteger draw per forwarded packet from the
    // switch's stream, in the switch's arrival order over all ports: at each step the earliest head
    // of the ports' arrival lists, equal times decided in port order by frame_before. Packets
    // arriving at a down port draw nothing; a packet with no route draws nothing and gets 0.
    // The k lanes of the switch run the merge in lockstep, each holding its own port's list: a
    // step exchanges the heads through LDS, every lane finds the same winner and makes the same
    // draw (each lane carries the stream), the winner records it. Heads and route counts of the
    // first 32 arrivals of a port sit in LDS (computed by the port's lane beforehand); frames are
    // read only to break equal times. The phases (init, then put / pick until done, then end) run
    // with wave_sync() between them; a CPU check may run each phase over the lanes in turn.
    static constexpr int SPRAY_CACHE = 32;
    static constexpr uint32_t SPRAY_NONE = 0xFFFFFFFFu;
    struct Spray
    {
        In in;
        int c, p;      // this port's arrivals, the next one's place
        uint32_t head; // its arrival time - slot start (SPRAY_NONE: none left)
        TwMrg g;       // the switch's stream (every lane the same)
        bool done;
    };
    // the packed entry of arrival j: time offset (20 bits), route count or 255 (a down port)
    static __device__ uint32_t spray_code(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t S0, const FrameHead &f)
    {
        const int ep = W.n_fail ? epoch_of(B, W, f.ts) : 0;
        const uint32_t nr = port_down(B, W, x.swl, x.q, ep) ? 255u : (uint32_t)routes_to(B, W, x.swl, f.dst, ep);
        return (uint32_t)(f.ts - S0) | nr << 20;
    }
    // this port's lane in the wavefront (its switch's first lane, base, plus the port)
    static __device__ int spray_lane(const SwLane &x, const SwLds &L) { return sw_lane_of(L, x); }
    static __device__ uint32_t spray_entry(const Batch &B, const WorldDesc &W, const SwLane &x, const SwLds &L, const Spray &sp, int64_t S0, int j)
    {
        if (j < SPRAY_CACHE)
            return (uint32_t)L.wave[spray_lane(x, L) * 33 + j];
        return spray_code(B, W, x, S0, sp.in.head(j));
    }
    static __device__ void spray_init(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t t, const SwLds &L, Spray &sp)
    {
        sp.c = 0;
        sp.p = 0;
        sp.head = SPRAY_NONE;
        sp.done = !x.active;
        if (!x.active)
            return;
        const int64_t S0 = t * SLOT_PS;
        sp.in = lane_input(B, W, x, t, L, x.q);
        sp.c = sp.in.count();
        const int lane = spray_lane(x, L);
        const int n = sp.c < SPRAY_CACHE ? sp.c : SPRAY_CACHE;
        for (int j = 0; j < n; j++)
            L.wave[lane * 33 + j] = (int)spray_code(B, W, x, S0, sp.in.head(j));
        if (sp.c > 0)
            sp.head = (uint32_t)L.wave[lane * 33]; // (the lane's own write)
        const TwMrgState &gs = B.sw_rng[W.rng0 + x.swl];
#pragma unroll
        for (int i = 0; i < 6; i++)
            sp.g.s[i] = gs.s[i];
    }
    static __device__ void spray_put(const SwLane &x, const SwLds &L, Spray &sp)
    {
        if (!x.active)
            return;
        L.cnt[x.q] = (int)sp.head;
        L.base[x.q] = sp.p;
    }
    // one step of the merge; true when the switch has no arrivals left
    static __device__ bool spray_pick(const Batch &B, const WorldDesc &W, WorldRun *R, const SwLane &x, int64_t t, const SwLds &L, Spray &sp)
    {
        if (!x.active || sp.done)
            return true;
        const int k = W.k;
        const int64_t S0 = t * SLOT_PS;
        // the earliest head (the first port holding it)
        int best = -1;
        uint32_t bt = SPRAY_NONE;
        for (int q = 0; q < k; q++)
        {
            const uint32_t h = (uint32_t)L.cnt[q] & 0xFFFFFu;
            if (L.cnt[q] != (int)SPRAY_NONE && (best < 0 || h < bt))
            {
                best = q;
                bt = h;
            }
        }
        if (best < 0)
        {
            sp.done = true;
            return true;
        }
        // equal times: port order, frame_before decides (the reference's scan over the ports)
        for (int q = best + 1; q < k; q++)
        {
            if (L.cnt[q] == (int)SPRAY_NONE || ((uint32_t)L.cnt[q] & 0xFFFFFu) != bt)
                continue;
            const Frame *f = lane_input(B, W, x, t, L, q).at(L.base[q]);
            const Frame *b = lane_input(B, W, x, t, L, best).at(L.base[best]);
            if (frame_before(R, f, b))
                best = q;
        }
        const uint32_t nr = (uint32_t)L.cnt[best] >> 20;
        uint8_t draw = 0;
        if (nr != 255u && nr)
            draw = (uint8_t)tw_mrg_integer(sp.g, nr);
        if (best == x.q)
        {
            if (nr != 255u)
                x.scr_up[x.q * 2 * MAX_TX + sp.p] = draw;
            sp.p++;
            sp.head = sp.p < sp.c ? spray_entry(B, W, x, L, sp, S0, sp.p) : SPRAY_NONE;
        }
        return false;
    }
    static __device__ void spray_end(const Batch &B, const WorldDesc &W, const SwLane &x, const SwLds &L, Spray &sp)
    {
        if (!x.active)
            return;
        if (x.q == 0)
        {
            TwMrgState &gs = B.sw_rng[W.rng0 + x.swl];
#pragma unroll
            for (int i = 0; i < 6; i++)
                gs.s[i] = sp.g.s[i];
        }
        // the count matrix row this lane's cache overlapped starts from zero again
        int *row = L.rows + x.q * (W.k + 1);
        for (int i = 0; i <= W.k; i++)
            row[i] = 0;
    }
};

} // namespace tw
