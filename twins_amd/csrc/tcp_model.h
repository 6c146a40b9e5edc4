// The TCP packet model on gfx950: one slot of a host (a lane runs its serial event loop) and one
// slot of a switch (a group of lanes of one wavefront, one lane per port), over the batch
// layout of tcp_state.h. The semantics follow TwiNS's model exactly (reference: src/step.cuh,
// src/kernel.cuh), so every flow's times equal sequential ns-3's; the comments cite the
// reference lines each part follows.
//
// Build with -ffp-contract=off: the only fused multiply-adds are the explicit fma() calls.
#pragma once
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
    static __device__ int chain_cmp(const uint32_t *a, const uint32_t *b)
    {
        int ia = 0, ib = 0;
        uint32_t ra = a[0] >> 16, rb = b[0] >> 16;
        for (;;)
        {
            uint32_t ca = a[ia] & 0xFFFF, cb = b[ib] & 0xFFFF;
            if (ca == 0 && cb != 0 && (a[RUNS] & 0xFFFF) != 0)
                return -1;
            if (cb == 0 && ca != 0 && (b[RUNS] & 0xFFFF) != 0)
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
                ra = a[ia] >> 16;
            }
            if (rb == 0)
            {
                if (++ib == RUNS)
                    break;
                rb = b[ib] >> 16;
            }
        }
        uint32_t xa = a[RUNS] << 16 | a[RUNS] >> 16, xb = b[RUNS] << 16 | b[RUNS] >> 16; // the root keys
        if (xa == 0 || xb == 0 || xa == xb)
            return 0;
        return xa < xb ? -1 : 1;
    }
    static __device__ __forceinline__ void count(WorldRun *R, int i, unsigned long long n = 1) { atomicAdd(&R->ev[i], n); }

    // -------------------------------------------------------------- egress FIFO
    // (src/step.cuh:167-317) frames not yet started; starts go to this slot's link message
    struct Fifo
    {
        const Batch *B;
        WorldRun *R;
        int lid;
        Frame *ring;
        int cap, head, size, red_k, limit, sent;
        int64_t link_free, last_start, woke, idle_start, slot_end;
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
            sent = 0;
            out = b.msg + ((size_t)ring_t * b.nlinks + link) * MAX_TX;
        }
        // the slot's message and the queue state back to memory
        __device__ __forceinline__ void store(int64_t t, int ring_t)
        {
            B->msg_cnt[(size_t)ring_t * B->nlinks + lid] = (long long)t << 16 | sent;
            B->q_head[lid] = head;
            B->q_size[lid] = size;
            B->q_link_free[lid] = link_free;
            B->q_last_start[lid] = last_start;
            chain_store(B->q_last_chain + (size_t)lid * CW, last);
            if (sent)
                count(R, EV_PKT_EVENTS, (unsigned long long)sent);
        }
        // start every frame whose transmission begins at or before `until` within the slot
        __device__ void advance(int64_t until)
        {
            while (size > 0 && sent < MAX_TX)
            {
                const Frame &f = ring[head];
                int64_t start = f.ts > link_free ? f.ts : link_free;
                if (start > until || start >= slot_end)
                    return;
                int64_t end = start + (int64_t)f.len * PS_PER_BYTE;
                Frame o = f;
                // the start event: the previous frame's TransmitComplete if this one waited,
                // else its own enqueue (a link that never sent is idle)
                if (f.ts <= link_free && link_free > last_start)
                {
                    chain_child(last, (uint32_t)((link_free - last_start) / PS_PER_BYTE), last);
                    chain_store(o.chain, last);
                }
                else
                {
                    idle_start = start;
                    chain_load(last, o.chain);
                }
                last_start = start;
                o.ts = end + PROP_PS;
                out[sent++] = o;
                link_free = end;
                head = head + 1 == cap ? 0 : head + 1;
                size--;
            }
            if (size > 0 && sent >= MAX_TX)
                count(R, EV_TXCAP);
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
        // RED (MinTh = MaxTh = K, QW = 1) on the queue-disc length: the frame in ns-3's
        // one-packet device queue is left out unless the Wake that refills it has yet to run
        __device__ bool enqueue(const Frame &f, int64_t now)
        {
            advance(now);
            bool refilled = last_start < now || (woke == now && idle_start == now);
            int nq = (size > 0 && link_free > now && refilled) ? size - 1 : size;
            Frame g = f;
            g.ts = now;
            if (nq >= red_k && nq > 1)
            {
                if (g.flags & F_ECT)
                {
                    g.flags |= F_CE;
                    count(R, EV_MARK);
                }
                else
                {
                    count(R, EV_DROP);
                    return false;
                }
            }
            if (nq >= limit)
            {
                count(R, EV_DROP);
                return false;
            }
            if (B->qt.rec && now >= B->qt.from && now < B->qt.until)
                trace(now, nq + 1, g.flow, g.sa, g.len, g.flags, g.to_sender);
            if (size >= cap)
            {
                count(R, EV_RING_OVF);
                return false;
            }
            int at = head + size;
            ring[at >= cap ? at - cap : at] = g;
            size++;
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
        int64_t slot_end;
        // the records of the connection side an event works on, copied into this lane's LDS
        // slots for the event (snd_open / rcv_open) and written back once at its end
        SndState *ls;
        RcvState *lr;
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
    static __device__ __forceinline__ SndState &snd_open(Host &x, int f)
    {
        rec_move(x.ls, &x.B->snd[f]);
        return *x.ls;
    }
    static __device__ __forceinline__ void snd_close(Host &x, int f) { rec_move(&x.B->snd[f], x.ls); }
    static __device__ __forceinline__ RcvState &rcv_open(Host &x, int f)
    {
        rec_move(x.lr, &x.B->rcv[f]);
        return *x.lr;
    }
    static __device__ __forceinline__ void rcv_close(Host &x, int f) { rec_move(&x.B->rcv[f], x.lr); }

    // Active lists: a host's sending and receiving connections. A connection leaves its list at
    // the end of a slot in which it is finished (sender: closed; receiver: the sender closed and
    // no receiver timer pending), as in the reference (src/step.cuh:2470-2510; with a full list
    // the add is refused and counted, so the removal's timing decides results in that regime);
    // compact() drops those and returns the earliest timer of the rest.
    static __device__ __forceinline__ int64_t compact_snd(const Batch &B, int h)
    {
        int64_t next = INT64_MAX;
        int32_t *l = B.act_snd + (size_t)h * MAX_ACTIVE;
        int k = 0, n = B.n_snd[h];
        for (int i = 0; i < n; i++)
        {
            int f = l[i];
            const SndState &s = B.snd[f];
            if (s.st == S_DONE)
                continue;
            l[k++] = f;
            if (s.rto_at >= 0 && s.rto_at < next)
                next = s.rto_at;
            if (s.syn_at >= 0 && s.syn_at < next)
                next = s.syn_at;
        }
        B.n_snd[h] = k;
        return next;
    }
    static __device__ __forceinline__ int64_t compact_rcv(const Batch &B, int h)
    {
        int64_t next = INT64_MAX;
        int32_t *l = B.act_rcv + (size_t)h * MAX_ACTIVE;
        int k = 0, n = B.n_rcv[h];
        for (int i = 0; i < n; i++)
        {
            int f = l[i];
            const RcvState &r = B.rcv[f];
            if (B.snd[f].st == S_DONE && r.delack_at < 0 && r.r_last_at < 0 && r.r_frx_at < 0)
                continue;
            l[k++] = f;
            int64_t v[4] = {r.delack_at, r.srv_syn_at, r.r_last_at, r.r_frx_at};
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (v[j] >= 0 && v[j] < next)
                    next = v[j];
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
    }
    static __device__ __forceinline__ void timer_chain(Host &x, uint32_t *base, int f, int64_t lead)
    {
        Chain c;
        chain_child(c, code_lead(lead), x.echain);
        chain_store(base + (size_t)f * CW, c);
    }
    static __device__ void emit(Host &x, int f, uint8_t flags, uint32_t seq, uint32_t ack, uint16_t len, bool to_sender, int64_t now)
    {
        const ConnStatic &cs = x.B->cs[f];
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
    static __device__ void send_ack(Host &x, int f, uint32_t ack, bool force_ece, bool no_ece, int64_t now)
    {
        RcvState &r = *x.lr;
        bool ece = false;
        if (force_ece)
            ece = true;
        else if (!no_ece && (r.ecn_r == E_CE_RCVD || r.ecn_r == E_SENDING_ECE))
        {
            ece = true;
            r.ecn_r = E_SENDING_ECE;
        }
        r.delack_at = -1;
        r.delack_count = 0;
        emit(x, f, F_ACK | (ece ? F_ECE : 0), 0, ack, CTRL_FRAME, true, now);
    }

    // -------------------------------------------------------------- the sent list
    // (src/step.cuh:409-975): TcpTxBuffer items (start, flags) in a ring of it_cap entries
    struct Tx
    {
        SndState &s;
        uint32_t *st;
        uint8_t *fl;
        uint32_t cap;
        __device__ Tx(const Batch &B, SndState &rec, int f) : s(rec)
        {
            const ConnStatic &c = B.cs[f];
            st = B.it_start + c.it_off;
            fl = B.it_fl + c.it_off;
            cap = c.it_cap;
        }
        __device__ uint32_t at(uint32_t i) const
        {
            uint32_t v = s.it_head + i;
            return v >= cap ? v - cap : v;
        }
        __device__ uint32_t beg(uint32_t i) const { return st[at(i)]; }
        __device__ uint32_t end(uint32_t i) const { return i + 1 < s.it_n ? st[at(i + 1)] : s.high; }
        __device__ uint8_t flags(uint32_t i) const { return fl[at(i)]; }
    };
    static __device__ __forceinline__ bool cand_of(uint8_t f) { return (f & (IT_LOST | IT_RETX | IT_SACKED)) == IT_LOST; }
    static __device__ void it_set(Tx &t, uint32_t i, uint8_t nf)
    {
        SndState &s = t.s;
        uint32_t k = t.at(i);
        uint8_t of = t.fl[k];
        if (of == nf)
            return;
        uint32_t len = t.end(i) - t.st[k];
        uint8_t d = of ^ nf;
        if (d & IT_SACKED)
        {
            s.sack_b += (nf & IT_SACKED) ? len : 0u - len;
            s.sacked_n += (nf & IT_SACKED) ? 1u : 0u - 1u;
        }
        if (d & IT_LOST)
            s.lost_b += (nf & IT_LOST) ? len : 0u - len;
        if (d & IT_RETX)
            s.rtx_b += (nf & IT_RETX) ? len : 0u - len;
        if (cand_of(of) != cand_of(nf))
        {
            if (cand_of(nf))
            {
                s.cand_n++;
                if (i < s.cand)
                    s.cand = i;
            }
            else
                s.cand_n--;
        }
        t.fl[k] = nf;
    }
    static __device__ bool it_insert(Tx &t, WorldRun *R, uint32_t i, uint32_t start, uint8_t f)
    {
        SndState &s = t.s;
        if (s.it_n >= t.cap)
        {
            count(R, EV_RTT_OVF);
            return false;
        }
        for (uint32_t j = s.it_n; j > i; j--)
        {
            uint32_t a = t.at(j), b = t.at(j - 1);
            t.st[a] = t.st[b];
            t.fl[a] = t.fl[b];
        }
        uint32_t k = t.at(i);
        t.st[k] = start;
        t.fl[k] = f;
        s.it_n++;
        if (s.cand >= i && s.cand > 0)
            s.cand++;
        if (cand_of(f))
        {
            s.cand_n++;
            if (i < s.cand)
                s.cand = i;
        }
        return true;
    }
    static __device__ void it_remove_start(Tx &t, uint32_t i)
    {
        SndState &s = t.s;
        for (uint32_t j = i; j + 1 < s.it_n; j++)
        {
            uint32_t a = t.at(j), b = t.at(j + 1);
            t.st[a] = t.st[b];
            t.fl[a] = t.fl[b];
        }
        s.it_n--;
        if (s.cand > i)
            s.cand--;
    }
    static __device__ uint32_t it_lower(const Tx &t, uint32_t seq)
    {
        uint32_t lo = 0, hi = t.s.it_n;
        while (lo < hi)
        {
            uint32_t m = (lo + hi) / 2;
            if (t.beg(m) < seq)
                lo = m + 1;
            else
                hi = m;
        }
        return lo;
    }
    static __device__ __forceinline__ uint32_t tb_sent(const SndState &s) { return s.it_n ? s.high - s.una : 0u; }
    static __device__ __forceinline__ uint32_t inflight(const SndState &s) { return tb_sent(s) - (s.sack_b + s.lost_b) + s.rtx_b; }
    static __device__ void tb_add_reno_sack(Tx &t)
    {
        SndState &s = t.s;
        uint32_t i = 1 + s.sacked_n;
        while (i < s.it_n && (t.flags(i) & IT_SACKED))
            i++;
        if (i < s.it_n)
        {
            it_set(t, i, t.flags(i) | IT_SACKED);
            s.sack_seen = 1;
            s.hisack = t.beg(i);
        }
    }
    static __device__ void tb_mark_head_lost(Tx &t)
    {
        if (t.s.it_n == 0)
            return;
        it_set(t, 0, (uint8_t)((t.flags(0) & ~(IT_SACKED | IT_RETX)) | IT_LOST));
    }
    static __device__ void tb_all_lost(Tx &t)
    {
        SndState &s = t.s;
        for (uint32_t i = 0; i < s.it_n; i++)
            t.fl[t.at(i)] = IT_LOST;
        s.sack_b = 0;
        s.sacked_n = 0;
        s.rtx_b = 0;
        s.lost_b = tb_sent(s);
        s.cand_n = s.it_n;
        s.cand = 0;
        s.sack_seen = 0;
        s.hisack = 0;
    }
    // DiscardUpTo(ack): the bytes of the items removed whole
    static __device__ uint32_t tb_discard(Tx &t, uint32_t ack)
    {
        SndState &s = t.s;
        uint32_t delivered = 0, removed = 0;
        while (s.it_n > 0)
        {
            uint32_t k = t.at(0), b = t.st[k], e = t.end(0);
            if (ack >= e)
            {
                if (t.fl[k])
                    it_set(t, 0, 0);
                s.it_head = t.at(1);
                s.it_n--;
                removed++;
                delivered += e - b;
            }
            else
            {
                if (ack > b)
                {
                    uint8_t fl = t.fl[k];
                    uint32_t off = ack - b;
                    if (fl & IT_SACKED)
                        s.sack_b -= off;
                    if (fl & IT_LOST)
                        s.lost_b -= off;
                    if (fl & IT_RETX)
                        s.rtx_b -= off;
                    t.st[k] = ack;
                }
                break;
            }
        }
        if (removed)
            s.cand = s.cand > removed ? s.cand - removed : 0;
        s.una = ack;
        if (s.it_n > 0 && (t.flags(0) & IT_SACKED))
        {
            it_set(t, 0, (uint8_t)(t.flags(0) & ~IT_SACKED));
            tb_add_reno_sack(t);
            tb_mark_head_lost(t);
        }
        if (s.hisack <= ack)
        {
            s.sack_seen = 0;
            s.hisack = 0;
        }
        return delivered;
    }
    static __device__ bool tb_next_seg(Tx &t, uint32_t *seq)
    {
        SndState &s = t.s;
        if (s.cand_n > 0)
        {
            for (uint32_t i = s.cand; i < s.it_n; i++)
            {
                if (!cand_of(t.flags(i)))
                    continue;
                s.cand = i;
                uint32_t b = t.beg(i);
                if (!s.sack_seen || b < s.hisack)
                {
                    *seq = b;
                    return true;
                }
                break;
            }
        }
        if (s.tail > s.high)
        {
            *seq = s.high;
            return true;
        }
        return false;
    }
    static __device__ bool tb_head_lost(Tx &t)
    {
        const SndState &s = t.s;
        if (!s.sack_seen || s.una >= s.hisack || s.it_n == 0)
            return false;
        return t.flags(0) & IT_LOST;
    }
    static __device__ bool tb_retx_ends_at(Tx &t, uint32_t x)
    {
        const SndState &s = t.s;
        if (s.it_n == 0 || x <= s.una || x > s.high)
            return false;
        uint32_t lo = it_lower(t, x);
        if (lo == 0)
            return false;
        uint32_t i = lo - 1;
        if (t.end(i) != x)
            return false;
        uint8_t fl = t.flags(i);
        return !(fl & IT_SACKED) && (fl & IT_RETX);
    }
    // GetTransmittedSegment for item i
    static __device__ uint32_t tb_retransmit(Tx &t, WorldRun *R, uint32_t i, uint32_t num)
    {
        SndState &s = t.s;
        uint32_t len = t.end(i) - t.beg(i), sz = num;
        if (i + 1 < s.it_n)
        {
            uint8_t a = t.flags(i), b = t.flags(i + 1);
            uint32_t nlen = t.end(i + 1) - t.beg(i + 1);
            if (!(b & IT_SACKED) && (a & IT_LOST) == (b & IT_LOST))
                sz = sz < len + nlen ? sz : len + nlen;
            else
                sz = sz < len ? sz : len;
        }
        else
            sz = sz < len ? sz : len;
        if (sz > len)
        {
            uint8_t a = t.flags(i), b = t.flags(i + 1);
            if ((a ^ b) & IT_RETX)
            {
                if (a & IT_RETX)
                    it_set(t, i, a &= (uint8_t)~IT_RETX);
                else
                    it_set(t, i + 1, b &= (uint8_t)~IT_RETX);
            }
            if (cand_of(b))
                s.cand_n--;
            it_remove_start(t, i + 1);
            len = t.end(i) - t.beg(i);
        }
        if (sz < len)
            it_insert(t, R, i + 1, t.beg(i) + sz, t.flags(i));
        uint8_t fl = t.flags(i);
        if (!(fl & IT_RETX))
            it_set(t, i, fl | IT_RETX);
        return sz;
    }
    // CopyFromSequence(num, seq)
    static __device__ uint32_t tb_copy(Tx &t, WorldRun *R, uint32_t num, uint32_t seq, bool *retx)
    {
        SndState &s = t.s;
        for (;;)
        {
            uint32_t avail = s.tail > seq ? s.tail - seq : 0;
            uint32_t sz = num < avail ? num : avail;
            if (sz == 0)
                return 0;
            uint32_t sent_end = s.high;
            if (sent_end >= seq + sz)
            {
                uint32_t i = it_lower(t, seq);
                if (i >= s.it_n || t.beg(i) != seq)
                {
                    count(R, EV_RTT_OVF);
                    return 0;
                }
                *retx = true;
                return tb_retransmit(t, R, i, sz);
            }
            if (sent_end <= seq)
            {
                if (!it_insert(t, R, s.it_n, seq, 0))
                    return 0;
                s.high = seq + sz;
                *retx = false;
                return sz;
            }
            num = sent_end - seq;
        }
    }

    // -------------------------------------------------------------- RTT (src/step.cuh:977-1126)
    static __device__ __forceinline__ uint32_t hat(const SndState &s, uint32_t cap, uint32_t j)
    {
        uint32_t v = s.h_head + j;
        return v >= cap ? v - cap : v;
    }
    static __device__ void rtt_history(Host &x, int f, uint32_t seq, uint32_t sz, bool retx, int64_t now)
    {
        const Batch &B = *x.B;
        WorldRun *R = x.R;
        SndState &s = *x.ls;
        const ConnStatic &c = B.cs[f];
        uint32_t *hs = B.h_seq + c.it_off, *hc = B.h_cnt + c.it_off;
        if (!retx)
        {
            if (s.h_n >= c.it_cap)
            {
                count(R, EV_RTT_OVF);
                return;
            }
            uint32_t k = hat(s, c.it_cap, s.h_n);
            hs[k] = seq;
            hc[k] = sz;
            B.h_t[c.it_off + k] = now;
            s.h_n++;
            return;
        }
        for (uint32_t j = 0; j < s.h_n; j++)
        {
            uint32_t k = hat(s, c.it_cap, j), n = hc[k] & 0x7FFFFFFFu;
            if (seq >= hs[k] && seq < hs[k] + n)
            {
                hc[k] = 0x80000000u | (seq + sz - hs[k]);
                return;
            }
        }
    }
    static __device__ void rtt_sample(SndState &s, int64_t m)
    {
        int64_t est = s.srtt, var = s.rttvar;
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
        s.srtt = est;
        s.rttvar = var;
        int64_t v = 4 * var;
        int64_t r = est + (v > CLOCK_GRAN_PS ? v : CLOCK_GRAN_PS);
        s.rto = r > MIN_RTO_PS ? r : MIN_RTO_PS;
    }
    static __device__ void rtt_ack(Host &x, int f, uint32_t ack, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        if (s.h_n == 0)
            return;
        const ConnStatic &c = B.cs[f];
        const uint32_t *hs = B.h_seq + c.it_off, *hc = B.h_cnt + c.it_off;
        uint32_t k = hat(s, c.it_cap, 0);
        if (!(hc[k] & 0x80000000u) && ack >= hs[k] + hc[k])
            rtt_sample(s, now - B.h_t[c.it_off + k]);
        while (s.h_n > 0)
        {
            k = hat(s, c.it_cap, 0);
            if (hs[k] + (hc[k] & 0x7FFFFFFFu) > ack)
                break;
            s.h_head = hat(s, c.it_cap, 1);
            s.h_n--;
        }
    }

    // -------------------------------------------------------------- congestion control
    // TcpCubic (src/step.cuh:1215-1360)
    static __device__ __forceinline__ uint32_t cub_u32(double d) { return (uint32_t)(uint64_t)(int64_t)d; }
    static __device__ void cub_hystart_reset(SndState &s, int64_t now)
    {
        s.cub_roundstart = s.cub_lastack = now;
        s.cub_endseq = s.st == S_EST ? s.high + 1 : 0;
        s.cub_currrtt = INT64_MIN;
        s.cub_samples = 0;
    }
    static __device__ __forceinline__ uint32_t cub_update(SndState &s, uint32_t segs, int64_t now)
    {
        const double C = 0.4, BETA = 0.7;
        uint32_t seg_cwnd = s.cwnd / MSS;
        s.cub_ackcnt += segs;
        if (s.cub_epoch == INT64_MIN)
        {
            s.cub_epoch = now;
            s.cub_ackcnt = segs;
            s.cub_tcpcwnd = seg_cwnd;
            if (s.cub_lastmax <= seg_cwnd)
            {
                s.cub_k = 0.0;
                s.cub_origin = seg_cwnd;
            }
            else
            {
                s.cub_k = tw_pow((double)(s.cub_lastmax - seg_cwnd) / C, 1 / 3.);
                s.cub_origin = s.cub_lastmax;
            }
        }
        int64_t t = (int64_t)((uint64_t)now + (uint64_t)s.cub_delaymin - (uint64_t)s.cub_epoch);
        double ts = tw_seconds(t), k = s.cub_k;
        double offs = ts < k ? k - ts : ts - k;
        uint32_t delta = cub_u32(C * tw_pow(offs, 3));
        uint32_t target = ts < k ? s.cub_origin - delta : s.cub_origin + delta;
        uint32_t cnt = target > seg_cwnd ? seg_cwnd / (target - seg_cwnd) : 100 * seg_cwnd;
        if (s.cub_lastmax == 0 && cnt > 20)
            cnt = 20;
        uint32_t scale = cub_u32(8 * (1024 + BETA * 1024) / 3 / (1024 - BETA * 1024));
        delta = (seg_cwnd * scale) >> 3;
        while (delta > 0 && s.cub_ackcnt > delta) // (delta >= 1 for any window of a segment or more)
        {
            s.cub_ackcnt -= delta;
            s.cub_tcpcwnd++;
        }
        if (s.cub_tcpcwnd > seg_cwnd)
        {
            uint32_t max_cnt = seg_cwnd / (s.cub_tcpcwnd - seg_cwnd);
            if (cnt > max_cnt)
                cnt = max_cnt;
        }
        return cnt > 2 ? cnt : 2;
    }
    static __device__ void cub_pkts_acked(SndState &s, int64_t now)
    {
        const int64_t ACK_DELTA = 2000000000, DELAY_MIN = 4000000000, DELAY_MAX = 1000000000000, CUBIC_DELTA = 10000000000;
        if (s.cub_epoch != INT64_MIN && now - s.cub_epoch < CUBIC_DELTA)
            return;
        int64_t rtt = s.srtt;
        if (s.cub_delaymin == INT64_MIN || s.cub_delaymin > rtt)
            s.cub_delaymin = rtt;
        uint32_t cwnd = s.cwnd;
        if (!(cwnd <= s.ssthresh && cwnd >= 16 * MSS) || s.cub_found)
            return;
        if (now - s.cub_lastack <= ACK_DELTA)
        {
            s.cub_lastack = now;
            if (now - s.cub_roundstart > s.cub_delaymin)
                s.cub_found = 1;
        }
        if (s.cub_samples < 8)
        {
            if (s.cub_currrtt == INT64_MIN || s.cub_currrtt > rtt)
                s.cub_currrtt = rtt;
            s.cub_samples++;
        }
        else
        {
            int64_t dm = s.cub_delaymin;
            int64_t th = dm > DELAY_MAX ? DELAY_MAX : (dm < DELAY_MIN ? DELAY_MIN : dm);
            if (s.cub_currrtt > dm + th)
                s.cub_found = 1;
        }
        if (s.cub_found)
            s.ssthresh = cwnd;
    }
    static __device__ void cub_increase(SndState &s, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (!s.cwnd_limited)
            return;
        if (s.cwnd < s.ssthresh)
        {
            if (ack + 1 > s.cub_endseq)
                cub_hystart_reset(s, now);
            s.cwnd += segs * MSS;
            segs = 0;
        }
        if (s.cwnd >= s.ssthresh && segs > 0)
        {
            s.cub_cnt += segs;
            uint32_t cnt = cub_update(s, segs, now);
            if (s.cub_cnt >= cnt)
            {
                s.cwnd += MSS;
                s.cub_cnt -= cnt;
            }
        }
    }
    static __device__ uint32_t cub_ssthresh(SndState &s)
    {
        const double BETA = 0.7;
        uint32_t seg_cwnd = s.cwnd / MSS;
        if (seg_cwnd < s.cub_lastmax)
            s.cub_lastmax = cub_u32((seg_cwnd * (1 + BETA)) / 2);
        else
            s.cub_lastmax = seg_cwnd;
        s.cub_epoch = INT64_MIN;
        uint32_t th = cub_u32(seg_cwnd * BETA);
        return (th > 2 ? th : 2) * MSS;
    }
    static __device__ void cub_loss(SndState &s, int64_t now)
    {
        s.cub_origin = 0;
        s.cub_k = 0;
        s.cub_ackcnt = 0;
        s.cub_tcpcwnd = 0;
        s.cub_delaymin = INT64_MIN;
        s.cub_found = 0;
        cub_hystart_reset(s, now);
    }
    // GetSsThresh (src/step.cuh:1057-1066)
    static __device__ uint32_t ssthresh_for(Host &x, SndState &s, uint32_t bif)
    {
        if (x.cubic)
            return cub_ssthresh(s);
        if (x.dctcp)
            return (uint32_t)((1.0 - s.alpha / 2.0) * s.cwnd);
        uint32_t half = bif / 2;
        return half > 2 * MSS ? half : 2 * MSS;
    }
    // PktsAcked (src/step.cuh:1362-1393)
    static __device__ void pkts_acked(Host &x, SndState &s, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (x.cubic)
        {
            cub_pkts_acked(s, now);
            return;
        }
        if (!x.dctcp)
            return;
        s.acked_total += segs * MSS;
        if (s.ecn_s == E_ECE_RCVD)
            s.acked_ecn += segs * MSS;
        if (!s.next_seq_flag)
        {
            s.next_seq = s.nxt;
            s.next_seq_flag = 1;
        }
        if (ack >= s.next_seq)
        {
            double m = s.acked_total > 0 ? (double)s.acked_ecn / s.acked_total : 0.0;
            s.alpha = __builtin_fma(1.0 - DCTCP_G, s.alpha, DCTCP_G * m);
            s.next_seq = s.nxt;
            s.acked_total = 0;
            s.acked_ecn = 0;
        }
    }
    // IncreaseWindow (src/step.cuh:1395-1455)
    static __device__ void increase_window(Host &x, SndState &s, uint32_t segs, uint32_t ack, int64_t now)
    {
        if (x.cubic)
        {
            cub_increase(s, segs, ack, now);
            return;
        }
        uint32_t cwnd = s.cwnd, ssth = s.ssthresh;
        if (x.dctcp)
        {
            if (cwnd < ssth)
            {
                if (segs >= 1)
                {
                    uint64_t n = (uint64_t)cwnd + (uint64_t)segs * MSS;
                    s.cwnd = (uint32_t)(n < ssth ? n : ssth);
                }
            }
            else
            {
                uint32_t w = cwnd / MSS;
                if (w == 0)
                    w = 1;
                uint32_t cnt = s.cwnd_cnt;
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
                s.cwnd = cwnd;
                s.cwnd_cnt = cnt;
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
        s.cwnd = cwnd;
    }

    // -------------------------------------------------------------- sending (src/step.cuh:1128-1213)
    static __device__ uint32_t send_segment(Host &x, int f, uint32_t seq, uint32_t max_size, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        Tx t(B, *x.ls, f);
        bool retx = false;
        uint32_t len = tb_copy(t, x.R, max_size, seq, &retx);
        if (len == 0)
            return 0;
        uint8_t flags = F_DATA;
        if (!retx && s.ecn_s == E_ECE_RCVD && s.ecn_echo_seq > s.ecn_cwr_seq)
        {
            s.ecn_s = E_CWR_SENT;
            s.ecn_cwr_seq = seq;
            flags |= F_CWR;
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
        rtt_history(x, f, seq, len, retx, now);
        if (seq + len > s.high)
            s.high = seq + len;
        atomicMax(&x.R->ev[EV_MAX_INFLIGHT], (unsigned long long)s.it_n);
        return len;
    }
    static __device__ void send_pending(Host &x, int f, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        if (s.st != S_EST)
            return;
        Tx t(B, *x.ls, f);
        for (int guard = 0; guard < 4096; guard++)
        {
            if (s.cong == C_OPEN && (s.fin & (FIN_SENT | FIN_ACKED)) == FIN_SENT)
                return;
            uint32_t pipe = inflight(s), cwnd = s.cwnd;
            uint32_t avail = cwnd > pipe ? cwnd - pipe : 0;
            if (avail == 0)
                return;
            uint32_t seq;
            if (!tb_next_seg(t, &seq))
                return;
            uint32_t remaining = s.tail - seq;
            if (avail < MSS && remaining > avail)
                return;
            s.nxt = seq;
            uint32_t len = send_segment(x, f, seq, avail < MSS ? avail : MSS, now);
            if (len == 0)
                return;
            s.nxt = seq + len;
        }
    }
    static __device__ void do_retransmit(Host &x, int f, int64_t now)
    {
        SndState &s = *x.ls;
        Tx t(*x.B, *x.ls, f);
        uint32_t seq;
        if (!tb_next_seg(t, &seq))
            seq = s.una;
        if (s.una >= s.tail)
            return;
        s.nxt = seq;
        send_segment(x, f, seq, MSS, now);
    }
    // NewAck (src/step.cuh:1457-1479)
    static __device__ void new_ack(Host &x, int f, uint32_t ack, bool reset_rto, int64_t now)
    {
        SndState &s = *x.ls;
        if (reset_rto)
        {
            int64_t est = s.srtt >= 0 ? s.srtt : INIT_RTT_PS, v = 4 * s.rttvar;
            int64_t r = est + (v > CLOCK_GRAN_PS ? v : CLOCK_GRAN_PS);
            s.rto = r > MIN_RTO_PS ? r : MIN_RTO_PS;
            s.rto_at = now + s.rto;
            timer_chain(x, x.B->rchain, f, s.rto);
        }
        if (ack >= s.tail && !(s.fin & FIN_SENT))
            s.rto_at = -1;
        if (ack > s.nxt)
            s.nxt = ack;
    }
    static __device__ __forceinline__ void enter_recovery(Host &x, int f, int64_t now)
    {
        SndState &s = *x.ls;
        Tx t(*x.B, *x.ls, f);
        tb_add_reno_sack(t);
        tb_mark_head_lost(t);
        s.recover = s.high;
        s.recover_active = 1;
        s.cong = C_RECOVERY;
        s.ssthresh = ssthresh_for(x, s, inflight(s) + MSS);
        s.cwnd = s.ssthresh;
        send_segment(x, f, s.una, MSS, now);
    }
    static __device__ void messages_acked(Host &x, int f, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        const ConnStatic &c = B.cs[f];
        while (s.m_dn < s.m_wr && s.una >= B.mg_end[c.m_first + s.m_dn])
        {
            B.mg_done[c.m_first + s.m_dn] = now;
            s.m_dn++;
            atomicAdd(&x.R->flows_done, 1u);
        }
        if (s.m_dn == c.m_cnt && (!c.fin_on || (s.fin & FIN_ACKED)))
        {
            s.st = S_DONE;
            s.rto_at = -1;
        }
    }
    // ReceivedAck + ProcessAck (src/step.cuh:1515-1692)
    static __device__ void on_ack(Host &x, int f, uint32_t ack, bool ece, int64_t now)
    {
        SndState &s = *x.ls;
        if (s.st != S_EST)
            return;
        uint32_t old = s.una;
        if (ack < old)
            return;
        Tx t(*x.B, *x.ls, f);
        rtt_ack(x, f, ack, now);
        bool advancing = ack > old;
        uint32_t delivered = 0;
        if (advancing)
            delivered = tb_discard(t, ack);
        if (s.cong == C_CWR && ack > s.recover)
        {
            s.cong = C_OPEN;
            s.cwnd = s.ssthresh;
        }
        if (advancing && x.dctcp && ece)
        {
            if (s.ecn_echo_seq < (int64_t)ack)
            {
                s.ecn_echo_seq = ack;
                s.ecn_s = E_ECE_RCVD;
                if (s.cong != C_CWR)
                {
                    s.ssthresh = ssthresh_for(x, s, inflight(s));
                    s.cwnd = s.ssthresh;
                    s.cong = C_CWR;
                    s.recover = s.high;
                }
            }
        }
        else if (s.ecn_s == E_ECE_RCVD && !ece)
            s.ecn_s = E_IDLE;
        bool dupack = ack == old && ack < s.high;
        uint32_t old_dup = s.dupacks;
        if (dupack)
        {
            uint8_t cg = s.cong;
            if (cg != C_LOSS)
            {
                if (cg != C_RECOVERY)
                {
                    s.dupacks = old_dup + 1;
                    old_dup = s.dupacks - 1;
                }
                if (cg == C_OPEN)
                    s.cong = cg = C_DISORDER;
                if (cg == C_RECOVERY)
                    tb_add_reno_sack(t);
                else if (cg == C_DISORDER)
                {
                    if (s.dupacks == (uint32_t)DUPTHRESH && (old >= s.recover || !s.recover_active))
                        enter_recovery(x, f, now);
                    else if (tb_head_lost(t))
                        enter_recovery(x, f, now);
                }
            }
        }
        if (ack == old && ack == s.high)
        {
        }
        else if (ack == old && ack > s.high)
            s.nxt = ack;
        else if (ack == old)
            pkts_acked(x, s, 1, ack, now);
        else if (advancing)
        {
            uint32_t segs = delivered / MSS;
            s.ack_rem += delivered % MSS;
            if (s.ack_rem >= MSS)
            {
                segs++;
                s.ack_rem -= MSS;
            }
            if (!dupack)
                s.dupacks = 0;
            uint8_t cg = s.cong;
            if (ack < s.recover && cg == C_RECOVERY)
            {
                tb_mark_head_lost(t);
                if (!tb_retx_ends_at(t, ack + MSS))
                    do_retransmit(x, f, now);
                pkts_acked(x, s, 1, ack, now);
                new_ack(x, f, ack, s.first_partial, now);
                s.first_partial = 0;
            }
            else if (ack < s.recover && cg == C_LOSS)
            {
                pkts_acked(x, s, segs, ack, now);
                increase_window(x, s, segs, ack, now);
                new_ack(x, f, ack, true, now);
            }
            else if (cg == C_CWR)
            {
                pkts_acked(x, s, segs, ack, now);
                new_ack(x, f, ack, true, now);
            }
            else
            {
                bool exited = false;
                if (cg == C_OPEN)
                    pkts_acked(x, s, segs, ack, now);
                else if (cg == C_DISORDER)
                {
                    if (segs >= old_dup)
                        pkts_acked(x, s, segs - old_dup, ack, now);
                    if (!dupack)
                        s.cong = C_OPEN;
                }
                else if (cg == C_RECOVERY)
                {
                    s.first_partial = 1;
                    segs = (ack - old) / MSS;
                    pkts_acked(x, s, segs, ack, now);
                    s.cong = C_OPEN;
                    exited = true;
                    s.dupacks = 0;
                }
                else if (cg == C_LOSS)
                {
                    s.first_partial = 1;
                    segs = (ack - s.recover) / MSS;
                    pkts_acked(x, s, segs, ack, now);
                    s.cong = C_OPEN;
                }
                if (ack >= s.recover)
                    s.recover_active = 0;
                if (exited)
                {
                    new_ack(x, f, ack, true, now);
                    s.cwnd = s.ssthresh;
                }
                if (s.cong == C_OPEN)
                {
                    increase_window(x, s, segs, ack, now);
                    new_ack(x, f, ack, true, now);
                }
            }
        }
        if ((s.fin & FIN_SENT) && s.it_n == 0 && ack == s.high + 1)
            s.fin |= FIN_ACKED;
        if (advancing)
        {
            messages_acked(x, f, now);
            if (s.st == S_DONE)
                return;
        }
        send_pending(x, f, now);
    }
    static __device__ __forceinline__ void on_rto(Host &x, int f, int64_t now)
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
        uint32_t before = inflight(s);
        s.dupacks = 0;
        Tx t(*x.B, *x.ls, f);
        tb_all_lost(t);
        s.recover = s.high;
        s.recover_active = 1;
        int64_t r = 2 * s.rto;
        s.rto = r < 60000000000000ll ? r : 60000000000000ll;
        s.h_n = 0;
        s.ssthresh = ssthresh_for(x, s, before);
        if (x.cubic)
            cub_loss(s, now);
        s.cong = C_LOSS;
        s.cwnd = MSS;
        send_pending(x, f, now);
    }

    // -------------------------------------------------------------- the receiving socket
    // (src/step.cuh:1727-2106)
    static __device__ void send_synack(Host &x, int f, int64_t now)
    {
        RcvState &r = *x.lr;
        int64_t rto = CONN_TIMEOUT_PS << r.srv_syn_n;
        r.srv_syn_n++;
        if (r.r_synack_t < 0)
            r.r_synack_t = now;
        emit(x, f, F_SYN | F_ACK, 0, 0, SYN_FRAME, true, now);
        if (r.srv_syn_at < 0)
        {
            r.srv_syn_at = now + rto;
            timer_chain(x, x.B->dchain, f, rto);
        }
    }
    static __device__ __forceinline__ void rcv_established(RcvState &r)
    {
        if (!r.r_est)
        {
            r.r_est = 1;
            r.srv_syn_at = -1;
        }
    }
    static __device__ void ce_state(Host &x, int f, bool ce_now, int64_t now)
    {
        RcvState &r = *x.lr;
        if (ce_now)
        {
            if (!r.ce_state && r.reserved && r.prior_flag)
                send_ack(x, f, r.prior_rcv_nxt, false, true, now);
            r.ce_state = 1;
            r.ecn_r = E_CE_RCVD;
        }
        else
        {
            if (r.ce_state && r.reserved && r.prior_flag)
                send_ack(x, f, r.prior_rcv_nxt, true, false, now);
            r.ce_state = 0;
            if (r.ecn_r == E_CE_RCVD || r.ecn_r == E_SENDING_ECE)
                r.ecn_r = E_IDLE;
        }
        r.prior_flag = 1;
        r.prior_rcv_nxt = r.rcv_nxt;
    }
    static __device__ void rx_ecn(Host &x, int f, const Frame &p, int64_t now)
    {
        if (!x.dctcp)
            return;
        RcvState &r = *x.lr;
        bool ce = p.flags & F_CE;
        if (ce && r.ecn_ce_seq < (int64_t)p.sa)
        {
            r.ecn_ce_seq = p.sa;
            r.ecn_r = E_CE_RCVD;
            ce_state(x, f, true, now);
        }
        else if (p.flags & F_ECT)
            ce_state(x, f, false, now);
    }
    static __device__ void snd_ecn(Host &x, SndState &s, const Frame &p)
    {
        if (!x.dctcp)
            return;
        if ((p.flags & F_CE) && !(s.s_ce & 1))
        {
            s.s_ce = 3;
            s.ecn_s = E_CE_RCVD;
        }
        else if (p.flags & F_ECT)
        {
            s.s_ce &= ~2;
            if (s.ecn_s == E_CE_RCVD || s.ecn_s == E_SENDING_ECE)
                s.ecn_s = E_IDLE;
        }
    }
    // TcpRxBuffer::Add for [seq, end): the out-of-order blocks stay disjoint and sorted
    static __device__ bool rx_add(Host &x, int f, uint32_t seq, uint32_t end)
    {
        const Batch &B = *x.B;
        WorldRun *R = x.R;
        RcvState &r = *x.lr;
        const ConnStatic &c = B.cs[f];
        uint32_t head = seq > r.rcv_nxt ? seq : r.rcv_nxt;
        if (head >= end)
            return false;
        uint32_t *bs = B.oo_s + c.oo_off, *be = B.oo_e + c.oo_off;
        const uint32_t cap = c.oo_cap, h0 = r.ooo_h;
        uint32_t n = r.ooo_n;
        auto at = [&](uint32_t i) { uint32_t v = h0 + i; return v >= cap ? v - cap : v; };
        uint32_t i;
        if (n == 0 || be[at(n - 1)] < head)
            i = n;
        else
        {
            uint32_t lo = 0, hi = n - 1;
            while (lo < hi)
            {
                uint32_t m = (lo + hi) / 2;
                if (be[at(m)] < head)
                    lo = m + 1;
                else
                    hi = m;
            }
            i = lo;
        }
        if (i < n && bs[at(i)] <= head && be[at(i)] >= end)
            return false;
        uint32_t us = head, ue = end, j = i;
        while (j < n && bs[at(j)] <= ue)
        {
            if (bs[at(j)] < us)
                us = bs[at(j)];
            if (be[at(j)] > ue)
                ue = be[at(j)];
            j++;
        }
        if (j == i)
        {
            if (n >= cap)
            {
                count(R, EV_OOO_OVF);
                return false;
            }
            for (uint32_t k = n; k > i; k--)
            {
                bs[at(k)] = bs[at(k - 1)];
                be[at(k)] = be[at(k - 1)];
            }
            n++;
        }
        else
        {
            uint32_t gone = j - i - 1;
            if (gone)
                for (uint32_t k = j; k < n; k++)
                {
                    bs[at(k - gone)] = bs[at(k)];
                    be[at(k - gone)] = be[at(k)];
                }
            n -= gone;
        }
        bs[at(i)] = us;
        be[at(i)] = ue;
        uint32_t h = h0;
        if (head == r.rcv_nxt)
        {
            r.rcv_nxt = be[at(0)];
            h = at(1);
            n--;
        }
        r.ooo_h = n ? h : 0;
        r.ooo_n = n;
        if (r.r_fin >= 0 && r.rcv_nxt == (uint32_t)r.r_fin)
            r.rcv_nxt++;
        return true;
    }
    static __device__ __forceinline__ bool rx_finished(const RcvState &r) { return r.r_fin >= 0 && (uint32_t)r.r_fin < r.rcv_nxt; }
    static __device__ void send_fin_ack(Host &x, int f, int64_t now)
    {
        RcvState &r = *x.lr;
        r.delack_at = -1;
        r.delack_count = 0;
        emit(x, f, F_ACK | F_FIN, 0, r.rcv_nxt, CTRL_FRAME, true, now);
        if (r.r_frx_at < 0)
        {
            int64_t est = r.r_srtt >= 0 ? r.r_srtt : INIT_RTT_PS, v = 4 * r.r_rttvar;
            int64_t rto = est + (v > CLOCK_GRAN_PS ? v : CLOCK_GRAN_PS);
            rto = rto > MIN_RTO_PS ? rto : MIN_RTO_PS;
            r.r_frx_at = now + rto;
            timer_chain(x, x.B->fchain, f, rto);
        }
    }
    static __device__ __forceinline__ int64_t last_rto(const RcvState &r)
    {
        int64_t est = r.r_srtt >= 0 ? r.r_srtt : INIT_RTT_PS, v = 4 * r.r_rttvar;
        return est + (v > CLOCK_GRAN_PS ? v : CLOCK_GRAN_PS);
    }
    static __device__ void do_peer_close(Host &x, int f, int64_t now)
    {
        RcvState &r = *x.lr;
        r.r_st = R_LAST_ACK;
        send_fin_ack(x, f, now);
        r.r_last_n = DATA_RETRIES;
        int64_t lr = last_rto(r);
        r.r_last_at = now + lr;
        timer_chain(x, x.B->dchain, f, lr);
    }
    static __device__ __forceinline__ void rcv_closed(RcvState &r)
    {
        r.r_st = R_CLOSED;
        r.r_last_at = -1;
        r.r_frx_at = -1;
        r.delack_at = -1;
    }
    static __device__ void received_data(Host &x, int f, uint32_t seq, uint32_t len, bool fin, int64_t now)
    {
        RcvState &r = *x.lr;
        uint32_t expected = r.rcv_nxt;
        if (!rx_add(x, f, seq, seq + len))
        {
            send_ack(x, f, r.rcv_nxt, false, false, now);
            return;
        }
        if (expected < r.rcv_nxt && rx_finished(r) && !fin)
        {
            do_peer_close(x, f, now);
            return;
        }
        bool gap = r.ooo_n > 0, filled = r.rcv_nxt > expected + len;
        if (gap || filled)
        {
            r.reserved = 0;
            send_ack(x, f, r.rcv_nxt, false, false, now);
            return;
        }
        uint32_t cnt = r.delack_count + 1u;
        if (cnt >= DELACK_COUNT)
        {
            r.reserved = 0;
            send_ack(x, f, r.rcv_nxt, false, false, now);
        }
        else
        {
            r.delack_count = (uint8_t)cnt;
            r.reserved = 1;
            if (r.delack_at < 0)
            {
                r.delack_at = now + DELACK_PS;
                timer_chain(x, x.B->dchain, f, DELACK_PS);
            }
        }
    }
    static __device__ void on_data(Host &x, int f, const Frame &p, int64_t now)
    {
        RcvState &r = *x.lr;
        if (!r.r_up || r.r_st == R_CLOSED)
            return;
        uint32_t seq = p.sa, len = (p.flags & F_DATA) ? p.len - 42u : 0u, end = seq + len;
        bool fin = p.flags & F_FIN;
        if (len > 0 && r.r_est && (r.r_st == R_LAST_ACK ? seq != r.rcv_nxt : end < r.rcv_nxt))
        {
            send_ack(x, f, r.rcv_nxt, false, true, now);
            return;
        }
        rx_ecn(x, f, p, now);
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
            received_data(x, f, seq, len, false, now);
            return;
        }
        if (seq < r.rcv_nxt)
            return;
        r.r_fin = end;
        if (r.rcv_nxt == end)
            r.rcv_nxt++;
        if (len)
            received_data(x, f, seq, len, true, now);
        if (!rx_finished(r))
            return;
        do_peer_close(x, f, now);
    }

    // -------------------------------------------------------------- application and dispatch
    static __device__ void add_active(Host &x, bool snd, int f)
    {
        int32_t *n = snd ? x.B->n_snd : x.B->n_rcv;
        int k = n[x.h];
        if (k >= MAX_ACTIVE)
        {
            count(x.R, EV_ACTIVE_OVF);
            return;
        }
        (snd ? x.B->act_snd : x.B->act_rcv)[(size_t)x.h * MAX_ACTIVE + k] = f;
        n[x.h] = k + 1;
    }
    // FlowSender::Write of the due messages; Socket::Send defers SendPendingData by 1 ps
    // (src/step.cuh:2120-2145)
    static __device__ void app_write(Host &x, int f, int64_t now)
    {
        const Batch &B = *x.B;
        SndState &s = *x.ls;
        const ConnStatic &c = B.cs[f];
        if (s.m_wr >= s.m_due)
            return;
        while (s.m_wr < s.m_due)
        {
            int m = c.m_first + s.m_wr;
            B.mg_est[m] = now;
            s.tail = B.mg_end[m];
            s.m_wr++;
        }
        if (c.fin_on && s.m_wr == c.m_cnt)
            s.fin |= FIN_CLOSE;
        uint32_t pipe = inflight(s);
        if (s.cwnd > pipe && s.syn_at < 0)
        {
            s.syn_at = now + 1;
            Chain ch;
            chain_child(ch, 2, x.echain);
            chain_store(B.schain + (size_t)f * CW, ch);
            due_insert(x, now + 1, f, 2);
        }
    }
    // a frame at its destination host (src/step.cuh:2147-2236): the sending side's part
    static __device__ void on_frame_snd(Host &x, int f, const Frame &p, int64_t now)
    {
        {
            SndState &s = *x.ls;
            if (p.flags & F_SYN)
            {
                if (s.st == S_SYN_SENT)
                {
                    s.st = S_EST;
                    s.syn_at = -1;
                    rtt_sample(s, now - s.timing_time);
                    emit(x, f, F_ACK, 0, 0, CTRL_FRAME, false, now);
                    chain_child(x.echain, 1, x.echain);
                    app_write(x, f, now);
                }
            }
            else if (p.flags & F_FIN)
            {
                if (s.fin & FIN_TW)
                    return;
                snd_ecn(x, s, p);
                on_ack(x, f, p.sa, p.flags & F_ECE, now);
                s.fin |= FIN_TW;
                emit(x, f, F_ACK, 0, 1, CTRL_FRAME, false, now);
            }
            else
            {
                if (s.fin & FIN_TW)
                    return;
                snd_ecn(x, s, p);
                on_ack(x, f, p.sa, p.flags & F_ECE, now);
            }
            return;
        }
    }
    // ... and the receiving side's
    static __device__ void on_frame_rcv(Host &x, int f, const Frame &p, int64_t now)
    {
        RcvState &r = *x.lr;
        if (!(p.flags & F_SYN) && r.r_up && r.r_srtt < 0 && r.r_synack_t >= 0)
        {
            int64_t m = now - r.r_synack_t;
            r.r_srtt = m;
            r.r_rttvar = m / 2;
        }
        if (p.flags & F_SYN)
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
            send_synack(x, f, now);
        }
        else if (p.flags & (F_DATA | F_FIN))
            on_data(x, f, p, now);
        else if (r.r_up && r.r_st != R_CLOSED)
        {
            rx_ecn(x, f, p, now);
            rcv_established(r);
            if (r.r_st == R_LAST_ACK)
                rcv_closed(r);
        }
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
    static __device__ __forceinline__ bool timer_before_arrival(Host &x, const Due &d, const Frame &p)
    {
        Chain arr, pc;
        chain_load(pc, p.chain);
        chain_child(arr, 2048u + p.len, pc);
        int r = chain_cmp(due_chain(*x.B, d), arr.r);
        if (r == 0)
            count(x.R, EV_UNDECIDED);
        return r <= 0;
    }

    // -------------------------------------------------------------- one slot of a host
    // (src/step.cuh:2263-2528): returns 1 if it sent or holds frames; *wake: its next slot
    static __device__ int host_step(const Batch &B, const WorldDesc &W, WorldRun *R, int hl, int64_t t, int64_t *wake, SndState *ls,
                                    RcvState *lr)
    {
        const int64_t S0 = t * SLOT_PS, E0 = S0 + SLOT_PS;
        const int h = W.host0 + hl, lid = W.link0 + hl;
        const int in = B.in_link[lid];
        const Frame *m2 = nullptr, *m1 = nullptr;
        int n2 = 0, n1 = 0, i2 = 0, i1 = 0;
        if (t >= 2)
        {
            size_t r = (size_t)((t - 2) % NRING) * B.nlinks + in;
            long long v = B.msg_cnt[r];
            n2 = (v >> 16) == t - 2 ? (int)(v & 0xFFFF) : 0;
            m2 = B.msg + r * MAX_TX;
            while (i2 < n2 && m2[i2].ts < S0)
                i2++;
        }
        if (t >= 1)
        {
            size_t r = (size_t)((t - 1) % NRING) * B.nlinks + in;
            long long v = B.msg_cnt[r];
            n1 = (v >> 16) == t - 1 ? (int)(v & 0xFFFF) : 0;
            m1 = B.msg + r * MAX_TX;
        }
        bool input = i2 < n2 || (n1 > 0 && m1[0].ts < E0);
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
        int pn = B.pend_next[h];
        const int pe = B.pend_end[h];
        for (;;)
        {
            int64_t tr = INT64_MAX, td = INT64_MAX, tp = INT64_MAX;
            if (i2 < n2)
                tr = m2[i2].ts;
            else if (i1 < n1 && m1[i1].ts < E0)
                tr = m1[i1].ts;
            if (x.id < x.nd)
                td = due[x.id].at;
            int pm = pn < pe ? B.pend_msg[pn] : -1;
            if (pm >= 0 && B.mg_start[pm] < E0)
                tp = B.mg_start[pm];
            int64_t now = tr < td ? tr : td;
            now = now < tp ? now : tp;
            if (now == INT64_MAX)
                break;
            if (now < S0)
                now = S0;
            x.nic.advance(now);
            if (now == tp && B.cs[B.mg_conn[pm]].m_first != pm)
            {
                // a later message of a connection (FlowSender::SendMessage)
                pn++;
                int f = B.mg_conn[pm];
                chain_root_msg(x.echain, pm - W.msg0);
                SndState &s = snd_open(x, f);
                s.m_due++;
                if (s.st == S_EST)
                    app_write(x, f, now);
                snd_close(x, f);
            }
            else if (now == tp)
            {
                // a connection's application starts: SYN
                pn++;
                int f = B.mg_conn[pm];
                SndState &s = snd_open(x, f);
                chain_root(x.echain, hl, f - W.conn0);
                s.m_due = 1;
                add_active(x, true, f);
                s.st = S_SYN_SENT;
                if (x.cubic)
                    cub_hystart_reset(s, now);
                s.rto = CONN_TIMEOUT_PS;
                s.syn_at = now + CONN_TIMEOUT_PS;
                timer_chain(x, B.rchain, f, CONN_TIMEOUT_PS);
                s.timing_time = now;
                emit(x, f, F_SYN, 0, 0, SYN_FRAME, false, now);
                snd_close(x, f);
            }
            else if (now == td && (now != tr || timer_before_arrival(x, due[x.id], i2 < n2 ? m2[i2] : m1[i1])))
            {
                Due d = due[x.id++];
                chain_load(x.echain, due_chain(B, d));
                const int f = d.flow;
                if (d.kind == 0)
                {
                    if (B.snd[f].rto_at == d.at)
                    {
                        snd_open(x, f);
                        on_rto(x, f, now);
                        snd_close(x, f);
                    }
                }
                else if (d.kind == 2)
                {
                    SndState &s = snd_open(x, f);
                    if (s.syn_at == d.at && s.st == S_EST)
                    {
                        s.syn_at = -1; // the deferred SendPendingData
                        send_pending(x, f, now);
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
                        send_ack(x, f, r.rcv_nxt, false, false, now);
                    }
                    else if (d.kind == 4 && r.r_last_at == d.at && r.r_st == R_LAST_ACK)
                    {
                        if (r.r_last_n == 0)
                            rcv_closed(r);
                        else
                        {
                            r.r_last_n--;
                            send_fin_ack(x, f, now);
                            int64_t lr = last_rto(r);
                            r.r_last_at = now + lr;
                            timer_chain(x, B.dchain, f, lr);
                        }
                    }
                    else if (d.kind == 5 && r.r_frx_at == d.at && r.r_st == R_LAST_ACK)
                    {
                        r.r_frx_at = -1;
                        send_fin_ack(x, f, now);
                    }
                    else if (d.kind == 3 && r.srv_syn_at == d.at && !r.r_est)
                    {
                        r.srv_syn_at = -1;
                        send_synack(x, f, now);
                    }
                    rcv_close(x, f);
                }
            }
            else
            {
                Frame p = i2 < n2 ? m2[i2++] : m1[i1++];
                Chain pc;
                chain_load(pc, p.chain);
                chain_child(x.echain, 2048u + p.len, pc);
                on_frame(x, p, now);
            }
        }
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
    static __device__ int route(const Batch &B, const WorldDesc &W, int sw, const Frame &p, int ep)
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
    static __device__ bool frame_before(WorldRun *R, const Frame &a, const Frame &b)
    {
        if (a.len != b.len)
            return a.len > b.len;
        int c = chain_cmp(a.chain, b.chain);
        if (c == 0)
            count(R, EV_UNDECIDED);
        return c != 0 ? c < 0 : a.flow < b.flow;
    }

    // -------------------------------------------------------------- one slot of a switch
    // (src/step.cuh:2706-2886). The k lanes of one switch within one wavefront, lane q = port q;
    // phases separated by wave_sync(). LDS per switch: rows[k][k+1] counts, in_win[k][3], base[k],
    // cnt[k]; per-wave global scratch holds the arrivals sorted per egress.
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
        __device__ const Frame &at(int j) const { return j < n2 - i2 ? m2[i2 + j] : m1[j - (n2 - i2)]; }
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
    static __device__ In lane_input(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t t, const SwLds &L, int q)
    {
        return input_of(B, B.in_link[W.link0 + W.H + x.swl * W.k + q], t, L.in_win + 3 * q);
    }
    static __device__ int lane_egress(const Batch &B, const WorldDesc &W, const SwLane &x, int j, const Frame &p)
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
        const int64_t S0 = t * SLOT_PS, E0 = S0 + SLOT_PS;
        const int in_link = B.in_link[x.lid];
        int w0 = 0, w1 = 0, w2 = 0;
        if (in_link >= 0)
        {
            if (t >= 2)
            {
                size_t r = (size_t)((t - 2) % NRING) * B.nlinks + in_link;
                long long v = B.msg_cnt[r];
                w1 = (v >> 16) == t - 2 ? (int)(v & 0xFFFF) : 0;
                const Frame *m = B.msg + r * MAX_TX;
                while (w0 < w1 && m[w0].ts < S0)
                    w0++;
            }
            if (t >= 1)
            {
                size_t r = (size_t)((t - 1) % NRING) * B.nlinks + in_link;
                long long v = B.msg_cnt[r];
                int n = (v >> 16) == t - 1 ? (int)(v & 0xFFFF) : 0;
                const Frame *m = B.msg + r * MAX_TX;
                while (w2 < n && m[w2].ts < E0)
                    w2++;
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
        const int n = in.count();
        for (int j = 0; j < n; j++)
        {
            int e = lane_egress(B, W, x, j, in.at(j));
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
    // 4: the start of egress q's region
    static __device__ void sw_base(const SwLane &x, const SwLds &L)
    {
        if (!x.active)
            return;
        int b = 0;
        for (int j = 0; j < x.q; j++)
            b += L.cnt[j];
        L.base[x.q] = b;
    }
    // 5: every arrival of ingress q into its egress's region, in arrival order
    static __device__ void sw_scatter(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t t, const SwLds &L)
    {
        if (!x.active)
            return;
        In in = lane_input(B, W, x, t, L, x.q);
        int *row = L.rows + x.q * (W.k + 1);
        const int n = in.count();
        for (int j = 0; j < n; j++)
        {
            const Frame &p = in.at(j);
            int e = lane_egress(B, W, x, j, p);
            if (e < 0)
                continue;
            int at = L.base[e] + row[e]++;
            x.scr_ts[at] = p.ts;
            x.scr_ref[at] = (uint32_t)(x.q << 16 | j);
        }
    }
    // 6: egress q's arrivals sorted by (time, frame_before), enqueued with the arrival event's
    // chain (RED), and the slot's transmissions; returns 1 if the port sent or holds frames
    static __device__ int sw_egress(const Batch &B, const WorldDesc &W, WorldRun *R, const SwLane &x, int64_t t, const SwLds &L)
    {
        if (!x.active)
            return 0;
        const int64_t E0 = (t + 1) * SLOT_PS;
        const int n = L.cnt[x.q], b0 = L.base[x.q];
        int64_t *ts = x.scr_ts + b0;
        uint32_t *ref = x.scr_ref + b0;
        auto frame_at = [&](uint32_t r) -> const Frame & { return lane_input(B, W, x, t, L, (int)(r >> 16)).at((int)(r & 0xFFFF)); };
        for (int i = 1; i < n; i++)
        {
            int64_t xv = ts[i];
            uint32_t r = ref[i];
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
        Fifo eg;
        const int rt = (int)(t % NRING);
        eg.load(B, R, x.lid, t, rt);
        for (int i = 0; i < n; i++)
        {
            Frame p = frame_at(ref[i]);
            Chain up, c;
            chain_load(up, p.chain);
            chain_child(c, 2048u + p.len, up);
            chain_store(p.chain, c);
            eg.enqueue(p, p.ts);
        }
        eg.advance(E0 - 1);
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
    static __device__ uint32_t spray_code(const Batch &B, const WorldDesc &W, const SwLane &x, int64_t S0, const Frame &f)
    {
        const int ep = W.n_fail ? epoch_of(B, W, f.ts) : 0;
        const uint32_t nr = port_down(B, W, x.swl, x.q, ep) ? 255u : (uint32_t)routes_to(B, W, x.swl, f.dst, ep);
        return (uint32_t)(f.ts - S0) | nr << 20;
    }
    // this port's lane in the wavefront (its switch's first lane, base, plus the port)
    static __device__ int spray_lane(const SwLane &x, const SwLds &L) { return (int)(L.cnt - L.wave - (64 * 33 + 64 * 3 + 64)) + x.q; }
    static __device__ uint32_t spray_entry(const Batch &B, const WorldDesc &W, const SwLane &x, const SwLds &L, const Spray &sp, int64_t S0, int j)
    {
        if (j < SPRAY_CACHE)
            return (uint32_t)L.wave[spray_lane(x, L) * 33 + j];
        return spray_code(B, W, x, S0, sp.in.at(j));
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
            L.wave[lane * 33 + j] = (int)spray_code(B, W, x, S0, sp.in.at(j));
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
            const Frame &f = lane_input(B, W, x, t, L, q).at(L.base[q]);
            const Frame &b = lane_input(B, W, x, t, L, best).at(L.base[best]);
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
