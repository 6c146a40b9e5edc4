// libtwins_core: the host side of the C ABI (include/twins.h). Builds worlds (topology, routes per
// failure epoch, connections and messages, ns-3 ECMP hashes, the per-host application order),
// packs batches of worlds into the device layout of tcp_state.h, and runs the slot kernels.
// Reference: the set-up and report parts of src/main.cu.
#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/twins.h"
#include "tcp_kernel.h"
#include "tcp_state.h"

using namespace tw;

// the batch record's place at the head of the static scratch (tw_batch_run_noalloc)
constexpr size_t NOALLOC_RECORD = 4096;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
static int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
extern "C" const char *tw_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------ ns-3 Hash32 (src/ns3hash.h:8-47)
// MurmurHash3_x86_32, seed 0x8BADF00D
static uint32_t hash32(const uint8_t *p, int n)
{
    auto rotl = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t h = 0x8BADF00Du;
    int i = 0;
    for (; i + 4 <= n; i += 4)
    {
        uint32_t k = (uint32_t)p[i] | (uint32_t)p[i + 1] << 8 | (uint32_t)p[i + 2] << 16 | (uint32_t)p[i + 3] << 24;
        h ^= rotl(k * c1, 15) * c2;
        h = rotl(h, 13) * 5 + 0xe6546b64u;
    }
    if (n > i)
    {
        uint32_t k = 0;
        for (int j = n - 1; j >= i; j--)
            k = (k << 8) | p[j];
        h ^= rotl(k * c1, 15) * c2;
    }
    h ^= (uint32_t)n;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
// the 5-tuple key ns-3's Ipv4GlobalRouting hashes: src IP, dst IP, protocol 6, ports (17 bytes)
static uint32_t flow_hash(uint32_t a, uint32_t b, uint16_t pa, uint16_t pb)
{
    uint8_t k[17] = {(uint8_t)(a >> 24), (uint8_t)(a >> 16), (uint8_t)(a >> 8), (uint8_t)a, (uint8_t)(b >> 24), (uint8_t)(b >> 16),
                     (uint8_t)(b >> 8), (uint8_t)b, 6, (uint8_t)(pa >> 8), (uint8_t)pa, (uint8_t)(pb >> 8), (uint8_t)pb};
    return hash32(k, 17);
}

#include "mrg_host.h"
#include "arena.h"

// ------------------------------------------------------------------ a world on the host
struct tw_world
{
    // topology
    int32_t kind = TOPO_FATTREE, k = 4, hk = 2, H = 0, S = 0, ecmp_n = 1, leaves = 0, spines = 0, hpl = 0, L = 0;
    std::vector<int32_t> host_sw, host_port;
    int32_t n_fail = 0;
    std::vector<int64_t> fail_t;
    std::vector<uint32_t> rt_off;
    std::vector<uint8_t> rt_n, rt_ports, port_dn;
    std::vector<int32_t> in_link; // [L] world-local
    std::vector<int32_t> hsub;    // per host: the subnet of its IP
    // flows: messages in file order, their new ids (each connection's consecutive)
    int32_t M = 0, NC = 0;
    std::vector<int32_t> f_src, f_dst;
    std::vector<int64_t> f_size, f_start;
    std::vector<int32_t> mnew;
    std::vector<int32_t> mg_conn;  // [new id]
    std::vector<uint32_t> mg_end;
    std::vector<int64_t> mg_start; // ps
    std::vector<int32_t> pend_msg, pend_next, pend_end;
    std::vector<ConnStatic> cs; // world-local ids and offsets
    uint64_t it_total = 0, oo_total = 0;
    // parameters
    int32_t cc = CC_DCTCP, ecn_k = RED_K, queue = QDISC_CAP, fin = 0;
    int64_t max_slots = 2000000;
    std::vector<TwMrgState> rng; // per switch (spray)
    // given ECMP hashes (--ecmp_from): connections whose computed hashes equal them, of hash_n
    int32_t hash_same = 0, hash_n = -1;
};

static int build_topology(tw_world &w, const tw_world_desc &d)
{
    if (d.topo == TW_TOPO_GRAPH)
    {
        const int nh = d.n_hosts, ns = d.n_switches;
        if (nh < 1 || ns < 1 || nh > 65535)
            return fail(TW_EINVAL, "topology: %d hosts, %d switches", nh, ns);
        w.kind = TOPO_GRAPH;
        w.H = nh;
        w.S = ns;
        w.ecmp_n = 1;
        std::vector<int> deg(ns, 0);
        std::vector<std::vector<std::pair<int, int>>> adj(ns); // (port, neighbour)
        w.host_sw.assign(nh, -1);
        w.host_port.assign(nh, 0);
        std::vector<int> hsub(nh);
        for (int i = 0; i < d.n_links; i++)
        {
            int a = d.link_a[i], b = d.link_b[i];
            if (b >= 0 || (a >= 0 && (a >= nh || w.host_sw[a] >= 0)) || -1 - b >= ns || (a < 0 && -1 - a >= ns))
                return fail(TW_EINVAL, "topology: hosts need exactly one link, as its first endpoint");
            int sb = -1 - b, pb = deg[sb]++;
            if (a >= 0)
            {
                w.host_sw[a] = sb;
                w.host_port[a] = pb;
                hsub[a] = i;
            }
            else
            {
                int sa = -1 - a, pa = deg[sa]++;
                adj[sa].push_back({pa, sb});
                adj[sb].push_back({pb, sa});
            }
        }
        for (int h = 0; h < nh; h++)
            if (w.host_sw[h] < 0)
                return fail(TW_EINVAL, "topology: host %d has no link", h);
        int k = 0;
        for (int x : deg)
            k = std::max(k, x);
        if (k > 32)
            return fail(TW_EINVAL, "topology: a switch has %d ports (at most 32)", k);
        w.k = k;
        w.hk = k / 2;
        // failures, in time order; epoch i follows the i-th (src/main.cu:373-464)
        struct Fl
        {
            int64_t t;
            int a, pa, b, pb;
            bool up;
        };
        std::vector<Fl> fails;
        for (int i = 0; i < d.n_fail; i++)
        {
            Fl f{d.fail_ns[i] * 1000, d.fail_a[i], -1, d.fail_b[i], -1, d.fail_up && d.fail_up[i]};
            if (f.a < 0 || f.a >= ns || f.b < 0 || f.b >= ns)
                return fail(TW_EINVAL, "--fail: no switch s%d or s%d", f.a, f.b);
            for (auto &pn : adj[f.a])
                if (pn.second == f.b && f.pa < 0)
                    f.pa = pn.first;
            for (auto &pn : adj[f.b])
                if (pn.second == f.a && f.pb < 0)
                    f.pb = pn.first;
            if (f.pa < 0)
                return fail(TW_EINVAL, "--fail: no link between s%d and s%d", f.a, f.b);
            fails.push_back(f);
        }
        std::stable_sort(fails.begin(), fails.end(), [](const Fl &x, const Fl &y) { return x.t < y.t; });
        const int NE = (int)fails.size() + 1;
        w.n_fail = NE - 1;
        for (auto &f : fails)
            w.fail_t.push_back(f.t);
        // equal-cost next hops per epoch: towards a neighbour one hop closer, in port order
        w.rt_off.assign((size_t)NE * ns * ns, 0);
        w.rt_n.assign((size_t)NE * ns * ns, 0);
        w.port_dn.assign((size_t)NE * ns * k, 0);
        std::vector<uint8_t> dn((size_t)ns * k, 0);
        std::vector<int> dist(ns), todo;
        for (int ep = 0; ep < NE; ep++)
        {
            if (ep > 0)
            {
                const Fl &f = fails[ep - 1];
                dn[(size_t)f.a * k + f.pa] = dn[(size_t)f.b * k + f.pb] = f.up ? 0 : 1;
            }
            std::copy(dn.begin(), dn.end(), w.port_dn.begin() + (size_t)ep * ns * k);
            for (int e = 0; e < ns; e++)
            {
                std::fill(dist.begin(), dist.end(), -1);
                todo.assign(1, e);
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
                        return fail(TW_EINVAL, "topology: switch %d cannot reach switch %d (epoch %d)", x, e, ep);
                    std::vector<int> ps;
                    for (auto &pn : adj[x])
                        if (!dn[(size_t)x * k + pn.first] && dist[pn.second] == dist[x] - 1)
                            ps.push_back(pn.first);
                    std::sort(ps.begin(), ps.end());
                    size_t r = ((size_t)ep * ns + x) * ns + e;
                    w.rt_off[r] = (uint32_t)w.rt_ports.size();
                    w.rt_n[r] = (uint8_t)ps.size();
                    for (int p : ps)
                        w.rt_ports.push_back((uint8_t)p);
                }
            }
        }
        // links: host h sends on link h, switch x port p on link H + x k + p
        w.L = nh + ns * k;
        w.in_link.assign(w.L, -1);
        auto sw_link = [&](int x, int p) { return nh + x * k + p; };
        for (int h = 0; h < nh; h++)
        {
            w.in_link[h] = sw_link(w.host_sw[h], w.host_port[h]);
            w.in_link[sw_link(w.host_sw[h], w.host_port[h])] = h;
        }
        for (int x = 0; x < ns; x++)
            for (auto &pn : adj[x])
            {
                // the neighbour's port on this link: the matching one of parallel links
                int y = pn.second, nth = 0;
                for (auto &q : adj[x])
                    if (q.second == y && q.first < pn.first)
                        nth++;
                for (auto &q : adj[y])
                    if (q.second == x && nth-- == 0)
                    {
                        w.in_link[sw_link(x, pn.first)] = sw_link(y, q.first);
                        break;
                    }
            }
        // ns-3's host IPs: 10.(n >> 8).(n & 255).1 with n the place of the host's link in the file
        w.hsub = hsub;
        return TW_OK;
    }
    if (d.n_fail)
        return fail(TW_EINVAL, "--fail needs --topo=graph");
    if (d.topo == TW_TOPO_LEAFSPINE)
    {
        const int L_ = d.leaves, S_ = d.spines, P = d.hpl;
        if (L_ < 1 || S_ < 1 || P < 1 || std::max(P + S_, L_) > 32 || L_ * P > 65535)
            return fail(TW_EINVAL, "leaf-spine: max(hpl + spines, leaves) <= 32");
        w.kind = TOPO_LEAFSPINE;
        w.k = std::max(P + S_, L_);
        w.hk = w.k / 2;
        w.leaves = L_;
        w.spines = S_;
        w.hpl = P;
        w.H = L_ * P;
        w.S = L_ + S_;
        w.ecmp_n = S_;
        w.L = w.H + w.S * w.k;
        w.in_link.assign(w.L, -1);
        auto sw_link = [&](int x, int p) { return w.H + x * w.k + p; };
        for (int h = 0; h < w.H; h++)
        {
            w.in_link[h] = sw_link(h / P, h % P);
            w.in_link[sw_link(h / P, h % P)] = h;
        }
        for (int l = 0; l < L_; l++)
            for (int j = 0; j < S_; j++)
            {
                w.in_link[sw_link(l, P + j)] = sw_link(L_ + j, l);
                w.in_link[sw_link(L_ + j, l)] = sw_link(l, P + j);
            }
    }
    else
    {
        const int k = d.k;
        if (k < 4 || k > 32 || (k & (k - 1)))
            return fail(TW_EINVAL, "FatTree: k must be a power of two in 4..32");
        w.kind = TOPO_FATTREE;
        w.k = k;
        w.hk = k / 2;
        w.H = k * k * k / 4;
        w.S = k * k + k * k / 4;
        w.ecmp_n = k / 2;
        w.L = w.H + w.S * k;
        const int hk = k / 2, E = k * hk;
        w.in_link.assign(w.L, -1);
        auto sw_link = [&](int x, int p) { return w.H + x * k + p; };
        for (int h = 0; h < w.H; h++)
        {
            w.in_link[h] = sw_link(h / hk, h % hk);
            w.in_link[sw_link(h / hk, h % hk)] = h;
        }
        for (int pod = 0; pod < k; pod++)
            for (int j = 0; j < hk; j++)
            {
                int agg = E + pod * hk + j;
                for (int x = 0; x < hk; x++)
                {
                    int edge = pod * hk + x;
                    w.in_link[sw_link(agg, x)] = sw_link(edge, j + hk);
                    w.in_link[sw_link(edge, j + hk)] = sw_link(agg, x);
                }
            }
        for (int i = 0; i < hk; i++)
            for (int j = 0; j < hk; j++)
            {
                int core = k * k + i * hk + j;
                for (int pod = 0; pod < k; pod++)
                {
                    int agg = E + pod * hk + i;
                    w.in_link[sw_link(core, pod)] = sw_link(agg, j + hk);
                    w.in_link[sw_link(agg, j + hk)] = sw_link(core, pod);
                }
            }
    }
    w.hsub.resize(w.H);
    for (int h = 0; h < w.H; h++)
        w.hsub[h] = h; // host subnets: the host index
    return TW_OK;
}

// messages over connections, the per-host application order, hashes and capacities
// (src/main.cu:531-806)
static int build_flows(tw_world &w, const tw_world_desc &d)
{
    const std::vector<int32_t> &hsub = w.hsub;
    const int M = d.n_flows, H = w.H;
    w.M = M;
    w.f_src.assign(d.src, d.src + M);
    w.f_dst.assign(d.dst, d.dst + M);
    w.f_size.assign(d.size, d.size + M);
    w.f_start.assign(d.start_ns, d.start_ns + M);
    std::vector<int> msg_conn(M);
    std::vector<std::vector<int>> conn_msgs;
    std::map<int64_t, int> by_id;
    for (int m = 0; m < M; m++)
    {
        if (w.f_src[m] < 0 || w.f_src[m] >= H || w.f_dst[m] < 0 || w.f_dst[m] >= H)
            return fail(TW_EINVAL, "flow %d: host outside the topology", m);
        if (w.f_size[m] < 0)
            return fail(TW_EINVAL, "flow %d: negative size", m);
        int64_t cid = d.conn ? d.conn[m] : -1;
        int c;
        if (cid < 0 || !by_id.count(cid))
        {
            c = (int)conn_msgs.size();
            conn_msgs.push_back({});
            if (cid >= 0)
                by_id[cid] = c;
        }
        else
        {
            c = by_id[cid];
            int first = conn_msgs[c][0], last = conn_msgs[c].back();
            if (w.f_src[first] != w.f_src[m] || w.f_dst[first] != w.f_dst[m] || w.f_start[m] < w.f_start[last])
                return fail(TW_EINVAL, "the flows of connection %lld need the same hosts and later starts", (long long)cid);
        }
        conn_msgs[c].push_back(m);
        msg_conn[m] = c;
    }
    const int NC = (int)conn_msgs.size();
    w.NC = NC;
    w.mnew.assign(M, 0);
    {
        int j = 0;
        for (int c = 0; c < NC; c++)
            for (int m : conn_msgs[c])
                w.mnew[m] = j++;
    }
    auto is_first = [&](int i) { return conn_msgs[msg_conn[i]][0] == i; };
    // per host, application events in ns-3's order at equal times: message writes (scheduled in
    // main) before connection starts (scheduled as the nodes initialize)
    std::vector<int> byold(M);
    for (int i = 0; i < M; i++)
        byold[i] = i;
    std::stable_sort(byold.begin(), byold.end(), [&](int a, int b) {
        if (w.f_src[a] != w.f_src[b])
            return w.f_src[a] < w.f_src[b];
        if (w.f_start[a] != w.f_start[b])
            return w.f_start[a] < w.f_start[b];
        bool fa = is_first(a), fb = is_first(b);
        if (fa != fb)
            return !fa;
        return fa ? msg_conn[a] < msg_conn[b] : a < b;
    });
    w.pend_msg.resize(M);
    for (int i = 0; i < M; i++)
        w.pend_msg[i] = w.mnew[byold[i]];
    w.pend_next.assign(H, 0);
    w.pend_end.assign(H, 0);
    for (int h = 0, i = 0; h < H; h++)
    {
        w.pend_next[h] = i;
        while (i < M && w.f_src[byold[i]] == h)
            i++;
        w.pend_end[h] = i;
    }
    w.mg_conn.assign(M, 0);
    w.mg_end.assign(M, 0);
    w.mg_start.assign(M, 0);
    std::vector<uint64_t> conn_bytes(NC, 0);
    for (int c = 0; c < NC; c++)
    {
        uint64_t off = 0;
        for (int m : conn_msgs[c])
        {
            off += (uint64_t)w.f_size[m];
            if (off >= 0xFFFFFFFFull)
                return fail(TW_EINVAL, "connection %d carries 4 GB or more", c);
            int j = w.mnew[m];
            w.mg_conn[j] = c;
            w.mg_end[j] = (uint32_t)off;
            w.mg_start[j] = w.f_start[m] * 1000;
        }
        conn_bytes[c] = off;
    }
    w.cs.assign(NC, ConnStatic{});
    for (int c = 0; c < NC; c++)
    {
        ConnStatic &r = w.cs[c];
        int first = conn_msgs[c][0];
        r.src = w.f_src[first];
        r.dst = w.f_dst[first];
        r.m_first = w.mnew[first];
        r.m_cnt = (int)conn_msgs[c].size();
        r.fin_on = d.fin != 0;
    }
    // ECMP hashes as ns-3 computes them: IPs 10.(n >> 8).(n & 255).1, sink port 1000, the
    // sender's ephemeral port from 49153 in its connections' (start, connection) order
    {
        auto ip = [&](int h) { return 0x0a000001u | (uint32_t)(hsub[h] & 0xFFFF) << 8; };
        std::vector<int> byhost(NC);
        for (int c = 0; c < NC; c++)
            byhost[c] = c;
        std::stable_sort(byhost.begin(), byhost.end(), [&](int a, int b) {
            int x = conn_msgs[a][0], y = conn_msgs[b][0];
            return w.f_src[x] != w.f_src[y] ? w.f_src[x] < w.f_src[y] : w.f_start[x] < w.f_start[y];
        });
        const bool given = d.hash_data && d.hash_ack;
        for (int i = 0, n = 0; i < NC; i++, n++)
        {
            int c = byhost[i];
            if (i > 0 && w.cs[byhost[i - 1]].src != w.cs[c].src)
                n = 0;
            if (n > 65535 - 49153 && !given)
                return fail(TW_EINVAL, "host %d binds more than %d connections: give the ECMP hashes", w.cs[c].src, 65535 - 49152);
            uint16_t sport = (uint16_t)(49153 + n), dport = 1000;
            uint32_t a = ip(w.cs[c].src), b = ip(w.cs[c].dst);
            w.cs[c].hash_data = flow_hash(a, b, sport, dport);
            w.cs[c].hash_ack = flow_hash(b, a, dport, sport);
        }
        if (given)
        {
            w.hash_same = w.hash_n = 0;
            for (int m = 0; m < M; m++)
                if (is_first(m))
                {
                    ConnStatic &r = w.cs[msg_conn[m]];
                    w.hash_n++;
                    w.hash_same += r.hash_data == d.hash_data[m] && r.hash_ack == d.hash_ack[m];
                    r.hash_data = d.hash_data[m];
                    r.hash_ack = d.hash_ack[m];
                }
        }
    }
    // the sent list / RTT history and the out-of-order blocks, sized per connection
    const uint64_t rtt_ring = d.rtt_ring > 0 ? (uint64_t)d.rtt_ring : 65536, ooo = d.ooo_segs > 0 ? (uint64_t)d.ooo_segs : 512;
    w.it_total = w.oo_total = 0;
    for (int c = 0; c < NC; c++)
    {
        uint64_t segs = (conn_bytes[c] + MSS - 1) / MSS + w.cs[c].m_cnt + 1;
        w.cs[c].it_cap = (uint32_t)std::min<uint64_t>(segs, rtt_ring);
        w.cs[c].oo_cap = (uint32_t)std::max(std::min<uint64_t>(segs / 2 + 2, ooo), std::min<uint64_t>((segs + 31) / 32, 1024));
        w.cs[c].it_off = w.it_total;
        w.cs[c].oo_off = w.oo_total;
        w.it_total += w.cs[c].it_cap;
        w.oo_total += w.cs[c].oo_cap;
    }
    return TW_OK;
}

extern "C" int tw_world_create(const tw_world_desc *d, tw_world **out)
{
    if (!d || !out)
        return fail(TW_EINVAL, "null argument");
    if (d->n_flows < 0 || (d->n_flows && (!d->src || !d->dst || !d->size || !d->start_ns)))
        return fail(TW_EINVAL, "flows: missing arrays");
    if (d->cc < 0 || d->cc > 2)
        return fail(TW_EINVAL, "unknown TCP");
    std::unique_ptr<tw_world> w(new tw_world);
    int rc = build_topology(*w, *d);
    if (rc)
        return rc;
    if ((rc = build_flows(*w, *d)))
        return rc;
    w->cc = d->cc;
    w->ecn_k = d->ecn_k > 0 ? d->ecn_k : RED_K;
    w->queue = d->queue > 0 ? d->queue : QDISC_CAP;
    w->fin = d->fin;
    w->max_slots = d->max_slots > 0 ? d->max_slots : 2000000;
    if (d->spray)
    {
        // node i's Ipv4GlobalRouting draws from stream i; switches are nodes H..
        w->rng.resize(w->S);
        for (int x = 0; x < w->S; x++)
            mrg_stream6(d->rng_seed ? d->rng_seed : 1, (uint64_t)(w->H + x), d->rng_run ? d->rng_run : 1, w->rng[x].s);
    }
    *out = w.release();
    return TW_OK;
}

extern "C" int tw_world_hash_check(const tw_world *w, int32_t *same, int32_t *conns)
{
    if (!w || !same || !conns)
        return fail(TW_EINVAL, "tw_world_hash_check: null argument");
    if (w->hash_n < 0)
        return fail(TW_EINVAL, "tw_world_hash_check: the world was made without given ECMP hashes (--ecmp_from)");
    *same = w->hash_same;
    *conns = w->hash_n;
    return TW_OK;
}

extern "C" int tw_world_dump_routes(const tw_world *w, const char *path)
{
    if (!w || !path)
        return fail(TW_EINVAL, "tw_world_dump_routes: null argument");
    if (w->kind != TOPO_GRAPH)
        return fail(TW_EINVAL, "tw_world_dump_routes: a topology-file world only (--topo=graph)");
    FILE *o = fopen(path, "w");
    if (!o)
        return fail(TW_EIO, "cannot write %s", path);
    // per switch and host: the ns-3 node (hosts first), the host, the output interfaces (port + 1)
    // of the equal-cost routes before any link failure (src/main.cu:466-486)
    const int ns = w->S;
    for (int x = 0; x < ns; x++)
        for (int h = 0; h < w->H; h++)
        {
            fprintf(o, "route %d %d", w->H + x, h);
            const int e = w->host_sw[h];
            if (e == x)
                fprintf(o, " %d", w->host_port[h] + 1);
            else
            {
                const size_t r = (size_t)x * ns + e;
                for (int i = 0; i < w->rt_n[r]; i++)
                    fprintf(o, " %d", w->rt_ports[w->rt_off[r] + i] + 1);
            }
            fprintf(o, "\n");
        }
    if (fclose(o))
        return fail(TW_EIO, "cannot write %s", path);
    return TW_OK;
}

extern "C" int tw_world_destroy(tw_world *w)
{
    delete w;
    return TW_OK;
}

static uint64_t world_bytes(const tw_world &w, int runs)
{
    const uint64_t F = runs == 25 ? 128 : 64, CHB = (runs + 1) * 4;
    uint64_t b = 0;
    b += (uint64_t)NRING * w.L * MAX_TX * F + (uint64_t)NRING * w.L * 8;          // link messages
    b += (uint64_t)w.L * (std::max(w.queue, 1) + 2) * F + (uint64_t)w.L * (40 + CHB); // queues
    b += (uint64_t)w.NC * (sizeof(ConnStatic) + sizeof(SndState) + sizeof(RcvState) + 4 * CHB);
    b += w.it_total * 21 + w.oo_total * 8;
    b += (uint64_t)w.M * 40 + (uint64_t)w.H * (2 * MAX_ACTIVE * 4 + MAX_DUE * 16 + 40);
    b += w.rt_off.size() * 5 + w.rt_ports.size() + w.port_dn.size() + w.rng.size() * 48;
    return b;
}

extern "C" int tw_world_info_get(const tw_world *w, tw_world_info *info)
{
    if (!w || !info)
        return fail(TW_EINVAL, "null argument");
    info->hosts = w->H;
    info->switches = w->S;
    info->ports_per_switch = w->k;
    info->links = w->L;
    info->conns = w->NC;
    info->flows = w->M;
    info->device_bytes = world_bytes(*w, 9);
    return TW_OK;
}

// ------------------------------------------------------------------ the `twins` argument set
static int read_flows(const std::string &path, int H, std::vector<int32_t> &src, std::vector<int32_t> &dst, std::vector<int64_t> &size,
                      std::vector<int64_t> &start, std::vector<int64_t> &conn)
{
    std::ifstream in(path);
    if (!in)
        return fail(TW_EIO, "cannot open %s", path.c_str());
    std::string line;
    while (std::getline(in, line))
    {
        if (line.empty() || !isdigit((unsigned char)line[0]))
            continue; // a header
        std::vector<int64_t> v;
        std::stringstream ss(line);
        for (std::string f; std::getline(ss, f, ',');)
            v.push_back(std::stoll(f));
        if (v.size() == 3) // node,size,start: the destination is node + H/2 (src/main.cu:129-130)
            v = {v[0], (v[0] + H / 2) % H, v[1], v[2], -1};
        else if (v.size() == 4)
            v.push_back(-1);
        else if (v.size() != 5)
            continue;
        src.push_back((int32_t)v[0]);
        dst.push_back((int32_t)v[1]);
        size.push_back(v[2]);
        start.push_back(v[3]);
        conn.push_back(v[4]);
    }
    return TW_OK;
}

extern "C" int tw_world_create_args(int argc, const char *const *argv, tw_world **out)
{
    if (!out)
        return fail(TW_EINVAL, "null argument");
    tw_world_desc d;
    memset(&d, 0, sizeof d);
    d.topo = TW_TOPO_FATTREE;
    d.k = 4;
    d.leaves = 16, d.spines = 8, d.hpl = 16;
    d.cc = TW_CC_DCTCP;
    d.rng_seed = 1;
    d.rng_run = 1;
    std::string flows, topo = "fattree", topofile, fail_spec, ecmp_from, tcp = "dctcp";
    try
    {
        for (int i = 0; i < argc; i++)
        {
            std::string a = argv[i];
            auto is = [&](const char *p) { return a.rfind(p, 0) == 0; };
            auto val = [&](const char *p) { return a.substr(strlen(p)); };
            if (is("--k="))
                d.k = std::stoi(val("--k="));
            else if (is("--flows="))
                flows = val("--flows=");
            else if (is("--tcp="))
                tcp = val("--tcp=");
            else if (is("--ecnK="))
                d.ecn_k = std::stoi(val("--ecnK="));
            else if (is("--queue="))
                d.queue = std::stoi(val("--queue="));
            else if (is("--fin="))
                d.fin = std::stoi(val("--fin="));
            else if (is("--topo="))
                topo = val("--topo=");
            else if (is("--topofile="))
                topofile = val("--topofile=");
            else if (is("--fail="))
                fail_spec = val("--fail=");
            else if (is("--leaves="))
                d.leaves = std::stoi(val("--leaves="));
            else if (is("--spines="))
                d.spines = std::stoi(val("--spines="));
            else if (is("--hpl="))
                d.hpl = std::stoi(val("--hpl="));
            else if (is("--spray="))
                d.spray = std::stoi(val("--spray="));
            else if (is("--RngSeed="))
                d.rng_seed = (uint32_t)std::stoll(val("--RngSeed="));
            else if (is("--RngRun="))
                d.rng_run = (uint64_t)std::stoll(val("--RngRun="));
            else if (is("--max_slots="))
                d.max_slots = std::stoll(val("--max_slots="));
            else if (is("--ecmp_from="))
                ecmp_from = val("--ecmp_from=");
            else if (is("--rtt_ring="))
                d.rtt_ring = std::stoi(val("--rtt_ring="));
            else if (is("--ooo_segs="))
                d.ooo_segs = std::stoi(val("--ooo_segs="));
            else
                return fail(TW_EINVAL, "unknown argument %s", a.c_str());
        }
    }
    catch (const std::exception &)
    {
        return fail(TW_EINVAL, "bad number in the arguments");
    }
    d.cc = tcp == "dctcp" ? TW_CC_DCTCP : tcp == "newreno" ? TW_CC_NEWRENO : tcp == "cubic" ? TW_CC_CUBIC : -1;
    if (d.cc < 0)
        return fail(TW_EINVAL, "--tcp: dctcp, newreno or cubic");
    if (flows.empty())
        return fail(TW_EINVAL, "--flows is required");
    std::vector<int32_t> la, lb, fa, fb;
    std::vector<int64_t> ft;
    std::vector<uint8_t> fu;
    int H = 0;
    if (topo == "graph")
    {
        d.topo = TW_TOPO_GRAPH;
        std::ifstream in(topofile);
        if (!in)
            return fail(TW_EIO, "cannot open %s", topofile.c_str());
        auto node = [](const std::string &x) { return x[0] == 'h' ? std::stoi(x.substr(1)) : -1 - std::stoi(x.substr(1)); };
        std::string key;
        try
        {
            while (in >> key)
            {
                if (key == "hosts")
                    in >> d.n_hosts;
                else if (key == "switches")
                    in >> d.n_switches;
                else if (key == "link")
                {
                    std::string a, b;
                    in >> a >> b;
                    la.push_back(node(a));
                    lb.push_back(node(b));
                }
                else
                    std::getline(in, key);
            }
        }
        catch (const std::exception &)
        {
            return fail(TW_EINVAL, "%s: bad link line", topofile.c_str());
        }
        d.n_links = (int32_t)la.size();
        d.link_a = la.data();
        d.link_b = lb.data();
        H = d.n_hosts;
        std::stringstream fs(fail_spec);
        for (std::string item; std::getline(fs, item, ',');)
        {
            if (item.empty())
                continue;
            std::vector<std::string> v;
            std::stringstream is(item);
            for (std::string x; std::getline(is, x, ':');)
                v.push_back(x);
            if (v.size() < 3 || v[1].size() < 2 || v[2].size() < 2)
                return fail(TW_EINVAL, "--fail: T:sA:sB[:up]");
            ft.push_back(std::stoll(v[0]));
            fa.push_back(std::stoi(v[1].substr(1)));
            fb.push_back(std::stoi(v[2].substr(1)));
            fu.push_back(v.size() > 3 && v[3] == "up");
        }
        d.n_fail = (int32_t)ft.size();
        d.fail_ns = ft.data();
        d.fail_a = fa.data();
        d.fail_b = fb.data();
        d.fail_up = fu.data();
    }
    else if (topo == "leafspine")
    {
        d.topo = TW_TOPO_LEAFSPINE;
        H = d.leaves * d.hpl;
        if (!fail_spec.empty())
            return fail(TW_EINVAL, "--fail needs --topo=graph");
    }
    else if (topo == "fattree")
    {
        H = d.k * d.k * d.k / 4;
        if (!fail_spec.empty())
            return fail(TW_EINVAL, "--fail needs --topo=graph");
    }
    else
        return fail(TW_EINVAL, "--topo: fattree, leafspine or graph");
    std::vector<int32_t> src, dst;
    std::vector<int64_t> size, start, conn;
    int rc = read_flows(flows, H, src, dst, size, start, conn);
    if (rc)
        return rc;
    d.n_flows = (int32_t)src.size();
    d.src = src.data();
    d.dst = dst.data();
    d.size = size.data();
    d.start_ns = start.data();
    d.conn = conn.data();
    std::vector<uint32_t> hd, ha;
    if (!ecmp_from.empty())
    {
        std::ifstream in(ecmp_from);
        if (!in)
            return fail(TW_EIO, "cannot open %s", ecmp_from.c_str());
        std::string line;
        std::getline(in, line);
        while ((int)hd.size() < d.n_flows && std::getline(in, line))
        {
            std::vector<std::string> v;
            std::stringstream ss(line);
            for (std::string f; std::getline(ss, f, ',');)
                v.push_back(f);
            if (v.size() < 8)
                return fail(TW_EINVAL, "%s has no hash columns", ecmp_from.c_str());
            hd.push_back((uint32_t)std::stoul(v[6]));
            ha.push_back((uint32_t)std::stoul(v[7]));
        }
        if ((int)hd.size() != d.n_flows)
            return fail(TW_EINVAL, "%s: %d rows for %d flows", ecmp_from.c_str(), (int)hd.size(), d.n_flows);
        d.hash_data = hd.data();
        d.hash_ack = ha.data();
    }
    return tw_world_create(&d, out);
}

// ------------------------------------------------------------------ batches
// A device-memory arena: one allocation made up front, carved first-fit in 256-byte units (the
// serving loop's requests: no hipMalloc or hipFree, which may wait for the device, while its
// persistent kernel runs).
struct tw_batch
{
    int runs = 9;
    std::vector<std::unique_ptr<tw_world>> worlds; // copies
    std::vector<void *> bufs; // every device allocation, freed with the batch
    void *batch_dev = nullptr; // BatchT<runs> in device memory
    std::vector<int32_t> msg0, conn0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    LaunchShape shape{};
    TeamState *teams = nullptr;
    int nteams = 0, nwq = 0;
    std::vector<TeamPlan> plan;
    int32_t *queue_next = nullptr;
    bool no_graph = false, graph_failed = false;
    Arena *arena = nullptr;     // allocations come from it (else hipMalloc)
    bool for_server = false;    // a serving-loop request (no launch plan of its own)
    hipStream_t up = nullptr;   // with an arena: uploads go on this stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    WorldRun *run = nullptr;
    std::vector<WorldRun> run_host;
    std::vector<int64_t> slots_before; // (per world, for the launch loops)
    std::vector<char> batch_host;      // the batch record as uploaded (BatchT<runs>)
    size_t off_teams = 0, off_queue_next = 0, off_scr_ts = 0, off_scr_ref = 0, off_scr_up = 0; // its scratch pointers
    int64_t *mg_est = nullptr, *mg_done = nullptr;
    QTrace qt{}; // the queue trace's buffer, when enabled
    ~tw_batch()
    {
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        if (graph)
            (void)hipGraphDestroy(graph);
        for (void *p : bufs)
            if (arena)
                arena->release(p);
            else
                (void)hipFree(p);
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

#define HIPCHK(x)                                                                                                           \
    do                                                                                                                      \
    {                                                                                                                       \
        hipError_t e_ = (x);                                                                                                \
        if (e_ != hipSuccess)                                                                                               \
            return fail(TW_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));                         \
    } while (0)

template <typename T>
static int dev_alloc(tw_batch &b, T **out, size_t n)
{
    size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
    void *p = nullptr;
    if (b.arena)
    {
        p = b.arena->alloc(bytes);
        if (!p)
            return fail(TW_ENOMEM, "the device arena has no %zu free bytes", bytes);
    }
    else
    {
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            return fail(TW_ENOMEM, "hipMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    b.bufs.push_back(p);
    *out = (T *)p;
    return TW_OK;
}
template <typename T>
static int dev_upload(tw_batch &b, T **out, const std::vector<T> &v)
{
    int rc = dev_alloc(b, out, v.size());
    if (rc)
        return rc;
    if (!v.empty())
    {
        if (b.up)
            HIPCHK(hipMemcpyAsync(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, b.up));
        else
            HIPCHK(hipMemcpy(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    }
    return TW_OK;
}
template <typename T>
static int dev_zero(tw_batch &b, T **out, size_t n)
{
    int rc = dev_alloc(b, out, n);
    if (rc)
        return rc;
    if (b.up)
        HIPCHK(hipMemsetAsync(*out, 0, sizeof(T) * std::max<size_t>(n, 1), b.up));
    else
        HIPCHK(hipMemset(*out, 0, sizeof(T) * std::max<size_t>(n, 1)));
    return TW_OK;
}

// Workgroups of a launch go to MI355X's 8 XCDs round-robin by workgroup id (XCD = id % 8), and
// each XCD has its own L2. Each rank of a team (largest teams first) takes the next id of the XCD
// with the most ids left, so a team's workgroups spread evenly over the XCDs and every team has
// the L2 capacity and bandwidth of all eight. Keeping each team on one XCD was measured slower:
// config 2 (three FatTree k=16 worlds, teams of 64) 234 ms against 184 ms per step, the config-4
// sweep 1.23-1.36 s against 0.99 s, a k=16 world alone 1.66-1.93 s against 0.91-1.03 s. Fills
// team_of (workgroup -> team << 8 | rank) and each team's lowest workgroup.
constexpr int XCDS = 8;
static void place_teams(std::vector<TeamPlan> &plan, std::vector<int32_t> &team_of)
{
    int n = 0;
    for (auto &p : plan)
        n += p.T;
    team_of.assign(n, -1);
    std::vector<std::vector<int>> freeb(XCDS);
    for (int b = n - 1; b >= 0; b--)
        freeb[b % XCDS].push_back(b); // popped from the back: lowest ids first
    std::vector<int> order(plan.size());
    for (size_t j = 0; j < plan.size(); j++)
        order[j] = (int)j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return plan[a].T > plan[c].T; });
    for (int j : order)
    {
        int rank = 0;
        plan[j].first = n;
        while (rank < plan[j].T)
        {
            int x = 0;
            for (int y = 1; y < XCDS; y++)
                if (freeb[y].size() > freeb[x].size())
                    x = y;
            const int b = freeb[x].back();
            freeb[x].pop_back();
            team_of[b] = j << 8 | rank++;
            plan[j].first = std::min(plan[j].first, b);
        }
    }
}

template <int RUNS>
static int pack(tw_batch &b, const tw_pack_cfg &cfg)
{
    typedef BatchT<RUNS> Batch;
    constexpr int CW = Batch::CW;
    Batch B;
    memset(&B, 0, sizeof B);
    const int W = (int)b.worlds.size();
    std::vector<WorldDesc> wd(W);
    int64_t H = 0, S = 0, L = 0, NC = 0, M = 0, IT = 0, OO = 0, RT = 0, RTP = 0, PDN = 0, FT = 0, HG = 0, RNG = 0;
    b.msg0.resize(W);
    b.conn0.resize(W);
    for (int i = 0; i < W; i++)
    {
        const tw_world &w = *b.worlds[i];
        WorldDesc &d = wd[i];
        memset(&d, 0, sizeof d);
        d.kind = w.kind, d.k = w.k, d.hk = w.hk, d.H = w.H, d.S = w.S, d.ecmp_n = w.ecmp_n;
        d.leaves = w.leaves, d.spines = w.spines, d.hpl = w.hpl, d.n_fail = w.n_fail;
        d.host0 = (int32_t)H, d.sw0 = (int32_t)S, d.link0 = (int32_t)L, d.L = w.L, d.conn0 = (int32_t)NC, d.nconn = w.NC;
        d.msg0 = (int32_t)M, d.M = w.M;
        d.rng0 = w.rng.empty() ? -1 : (int32_t)RNG;
        d.max_slots = w.max_slots;
        d.o_host = HG, d.o_rt = RT, d.o_rtp = RTP, d.o_pdn = PDN, d.o_fail = FT;
        b.msg0[i] = (int32_t)M;
        b.conn0[i] = (int32_t)NC;
        H += w.H, S += w.S, L += w.L, NC += w.NC, M += w.M, IT += w.it_total, OO += w.oo_total;
        RT += w.rt_off.size(), RTP += w.rt_ports.size(), PDN += w.port_dn.size(), FT += w.fail_t.size();
        HG += w.kind == TOPO_GRAPH ? w.H : 0;
        RNG += w.rng.size();
    }
    if (H > INT32_MAX / 2 || L > INT32_MAX / 2 || M > INT32_MAX / 2 || NC > INT32_MAX / 2)
        return fail(TW_EINVAL, "batch too large");
    B.nworlds = W, B.nhosts = (int32_t)H, B.nswitches = (int32_t)S, B.nlinks = (int32_t)L, B.nconns = (int32_t)NC, B.nmsgs = (int32_t)M;
    // host-side images
    std::vector<ConnStatic> cs(NC);
    std::vector<SndState> snd(NC);
    std::vector<RcvState> rcv(NC);
    std::vector<int32_t> mg_conn(M), pend_msg(M), host_world(H), pend_next(H), pend_end(H), in_link(L, -1), sw_world(S);
    std::vector<uint32_t> mg_end(M), g_rt_off(RT);
    std::vector<int64_t> mg_start(M), neg(M, -1), g_fail(FT), wake(H, 0), neg_links(L, 0);
    std::vector<uint8_t> host_cc(H), g_rt_n(RT), g_rt_ports(RTP), g_pdn(PDN);
    std::vector<int32_t> g_host_sw(HG), g_host_port(HG);
    std::vector<uint64_t> q_base(L);
    std::vector<int32_t> q_cap(L), q_red_k(L), q_limit(L);
    std::vector<TwMrgState> rng(RNG);
    std::vector<WorldRun> run(W);
    SndState s0;
    memset(&s0, 0, sizeof s0);
    s0.cwnd = INIT_CWND, s0.ssthresh = 0xFFFFFFFFu, s0.first_partial = 1;
    s0.cub_delaymin = s0.cub_epoch = s0.cub_roundstart = s0.cub_lastack = s0.cub_currrtt = INT64_MIN;
    s0.ecn_echo_seq = s0.ecn_cwr_seq = -1, s0.ecn_s = E_IDLE, s0.alpha = 1.0;
    s0.srtt = -1, s0.rttvar = 0, s0.rto = MIN_RTO_PS, s0.rto_at = -1, s0.timing_time = 0, s0.syn_at = -1;
    RcvState r0;
    memset(&r0, 0, sizeof r0);
    r0.r_st = R_OPEN, r0.r_fin = r0.r_synack_t = r0.r_srtt = r0.r_last_at = r0.r_frx_at = r0.srv_syn_at = r0.delack_at = r0.ecn_ce_seq = -1;
    r0.ecn_r = E_IDLE;
    uint64_t qf = 0;
    for (int i = 0; i < W; i++)
    {
        const tw_world &w = *b.worlds[i];
        const WorldDesc &d = wd[i];
        const uint64_t it0 = 0;
        (void)it0;
        int64_t it_base = 0, oo_base = 0;
        for (int j = 0; j < i; j++)
            it_base += b.worlds[j]->it_total, oo_base += b.worlds[j]->oo_total;
        for (int c = 0; c < w.NC; c++)
        {
            ConnStatic x = w.cs[c];
            x.src += d.host0;
            x.dst += d.host0;
            x.m_first += d.msg0;
            x.it_off += it_base;
            x.oo_off += oo_base;
            x.world = i;
            cs[d.conn0 + c] = x;
            snd[d.conn0 + c] = s0;
            rcv[d.conn0 + c] = r0;
        }
        for (int m = 0; m < w.M; m++)
        {
            mg_conn[d.msg0 + m] = w.mg_conn[m] + d.conn0;
            mg_end[d.msg0 + m] = w.mg_end[m];
            mg_start[d.msg0 + m] = w.mg_start[m];
            pend_msg[d.msg0 + m] = w.pend_msg[m] + d.msg0;
        }
        for (int h = 0; h < w.H; h++)
        {
            host_world[d.host0 + h] = i;
            pend_next[d.host0 + h] = w.pend_next[h] + d.msg0;
            pend_end[d.host0 + h] = w.pend_end[h] + d.msg0;
            host_cc[d.host0 + h] = (uint8_t)w.cc;
        }
        for (int x = 0; x < w.S; x++)
            sw_world[d.sw0 + x] = i;
        for (int l = 0; l < w.L; l++)
        {
            in_link[d.link0 + l] = w.in_link[l] < 0 ? -1 : w.in_link[l] + d.link0;
            q_base[d.link0 + l] = qf;
            q_cap[d.link0 + l] = w.queue + 2;
            q_red_k[d.link0 + l] = w.ecn_k;
            q_limit[d.link0 + l] = w.queue;
            qf += w.queue + 2;
        }
        if (w.kind == TOPO_GRAPH)
        {
            std::copy(w.host_sw.begin(), w.host_sw.end(), g_host_sw.begin() + d.o_host);
            std::copy(w.host_port.begin(), w.host_port.end(), g_host_port.begin() + d.o_host);
            std::copy(w.rt_off.begin(), w.rt_off.end(), g_rt_off.begin() + d.o_rt);
            std::copy(w.rt_n.begin(), w.rt_n.end(), g_rt_n.begin() + d.o_rt);
            std::copy(w.rt_ports.begin(), w.rt_ports.end(), g_rt_ports.begin() + d.o_rtp);
            std::copy(w.port_dn.begin(), w.port_dn.end(), g_pdn.begin() + d.o_pdn);
            std::copy(w.fail_t.begin(), w.fail_t.end(), g_fail.begin() + d.o_fail);
        }
        if (d.rng0 >= 0)
            std::copy(w.rng.begin(), w.rng.end(), rng.begin() + d.rng0);
        WorldRun &r = run[i];
        memset(&r, 0, sizeof r);
        r.act_cur = 1; // no idle skip before the first slots have run
        r.act_prev = 1;
        r.finished = w.M == 0;
    }
    int rc = 0;
#define UP(field, vec)                                                                                                      \
    if ((rc = dev_upload(b, &B.field, vec)))                                                                                \
        return rc;
#define ZERO(field, n)                                                                                                      \
    if ((rc = dev_zero(b, &B.field, n)))                                                                                    \
        return rc;
    UP(worlds, wd);
    UP(run, run);
    UP(cs, cs);
    UP(snd, snd);
    UP(rcv, rcv);
    ZERO(rchain, NC * CW);
    ZERO(schain, NC * CW);
    ZERO(dchain, NC * CW);
    ZERO(fchain, NC * CW);
    ZERO(it_start, IT);
    ZERO(h_seq, IT);
    ZERO(h_cnt, IT);
    ZERO(it_fl, IT);
    ZERO(h_t, IT);
    ZERO(oo_s, OO);
    ZERO(oo_e, OO);
    UP(mg_conn, mg_conn);
    UP(mg_end, mg_end);
    UP(mg_start, mg_start);
    UP(mg_est, neg);
    UP(mg_done, neg);
    UP(pend_msg, pend_msg);
    UP(host_world, host_world);
    UP(pend_next, pend_next);
    UP(pend_end, pend_end);
    ZERO(n_snd, H);
    ZERO(n_rcv, H);
    ZERO(act_snd, H * MAX_ACTIVE);
    ZERO(act_rcv, H * MAX_ACTIVE);
    UP(host_cc, host_cc);
    UP(wake, wake);
    {
        std::vector<int64_t> none(H, INT64_MAX);
        UP(tmin_s, none);
        UP(tmin_r, none);
    }
    ZERO(due, H * MAX_DUE);
    UP(in_link, in_link);
    UP(sw_world, sw_world);
    ZERO(qring, qf);
    UP(q_base, q_base);
    UP(q_cap, q_cap);
    ZERO(q_head, L);
    ZERO(q_size, L);
    UP(q_red_k, q_red_k);
    UP(q_limit, q_limit);
    ZERO(q_link_free, L);
    ZERO(q_last_start, L);
    ZERO(q_last_chain, L * CW);
    if ((rc = dev_alloc(b, &B.msg, (size_t)NRING * L * MAX_TX)))
        return rc;
    {
        std::vector<long long> init((size_t)NRING * L, msg_word(-1, 0, 0)); // of no slot
        UP(msg_cnt, init);
    }
    UP(g_host_sw, g_host_sw);
    UP(g_host_port, g_host_port);
    UP(g_rt_off, g_rt_off);
    UP(g_rt_n, g_rt_n);
    UP(g_rt_ports, g_rt_ports);
    UP(g_port_dn, g_pdn);
    UP(g_fail_t, g_fail);
    UP(sw_rng, rng);
    if (b.for_server)
    {
        // a serving-loop request: the server's kernel brings the teams and the scratch
        b.run = B.run;
        b.mg_est = B.mg_est;
        b.mg_done = B.mg_done;
        b.run_host = run;
        Batch *bd = nullptr;
        if ((rc = dev_alloc(b, &bd, 1)))
            return rc;
        HIPCHK(hipMemcpyAsync(bd, &B, sizeof B, hipMemcpyHostToDevice, b.up));
        HIPCHK(hipStreamSynchronize(b.up));
        b.batch_dev = bd;
        return TW_OK;
    }
    // launch plan: teams of workgroups, each taking worlds from one queue
    int dev = 0, cus = 256, per_cu = 1;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LaunchShape &sh = b.shape;
    // workgroup size: 256 threads (4 waves), the kernel's bound: it may then use 512 VGPRs and
    // runs fully inlined without spills; the GPU holds one such workgroup per CU, so worlds do not
    // share a CU (co-resident workgroups of other worlds slowed each other more than they hid each
    // other's latency)
    sh.threads = cfg.block_threads > 0 ? cfg.block_threads : 256;
    if (sh.threads % 64 || sh.threads > 256)
        return fail(TW_EINVAL, "block_threads: a multiple of 64 up to 256");
    HIPCHK(tcp_max_blocks(RUNS, sh.threads, &per_cu));
    const int resident = std::max(1, per_cu) * cus;
    auto size_of = [&](int i) { return (int64_t)b.worlds[i]->H + b.worlds[i]->S; };
    std::vector<int32_t> order(W), team_T, team_q, world_q(W, 0);
    int nq = 1;
    for (int i = 0; i < W; i++)
        order[i] = i;
    if (cfg.plan)
    {
        const tw_plan &p = *cfg.plan;
        if (p.nteams < 1 || p.nqueues < 1 || !p.team_blocks || !p.team_queue || !p.world_queue)
            return fail(TW_EINVAL, "plan: teams, queues and the world queues are required");
        nq = p.nqueues;
        team_T.assign(p.team_blocks, p.team_blocks + p.nteams);
        team_q.assign(p.team_queue, p.team_queue + p.nteams);
        world_q.assign(p.world_queue, p.world_queue + W);
        for (int j = 0; j < p.nteams; j++)
            if (team_T[j] < 1 || team_q[j] < 0 || team_q[j] >= nq)
                return fail(TW_EINVAL, "plan: team %d has %d workgroups, queue %d", j, team_T[j], team_q[j]);
        for (int i = 0; i < W; i++)
            if (world_q[i] < 0 || world_q[i] >= nq)
                return fail(TW_EINVAL, "plan: world %d in queue %d of %d", i, world_q[i], nq);
        if (p.world_order)
        {
            std::vector<char> seen(W, 0);
            for (int i = 0; i < W; i++)
            {
                int x = p.world_order[i];
                if (x < 0 || x >= W || seen[x])
                    return fail(TW_EINVAL, "plan: world_order is not a permutation of the worlds");
                seen[x] = 1;
                order[i] = x;
            }
        }
    }
    else
    {
        // one queue, largest worlds first (the long ones start early); few large worlds are
        // spread over the GPU, many run a workgroup each
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return size_of(x) > size_of(y); });
        // a team with a wave for every work item of the largest world, at one host per item for a
        // world alone (at most 128 workgroups) and two for each of several (at most 48). Measured
        // (slot barrier without cache fences): a FatTree k=16 world alone (k16_a2a50s, DCTCP) ran
        // 0.69, 0.57, 0.52, 0.49 and 0.47 s on teams of 32, 48, 64, 96 and 128; k8_paper (DCTCP)
        // 0.342, 0.344, 0.321 and 0.318 s on 26, 40, 52 and 64 (52: one host per item), k8_a2a30
        // 0.397, 0.397, 0.360 and 0.362 s; three k=16 worlds side by side (config 2) 134, 108, 100,
        // 97 and 98 ms per run on 16, 24, 32, 48 and 64.
        const tw_world &big = *b.worlds[order[0]];
        const int waves_per = std::max(1, (cfg.block_threads > 0 ? cfg.block_threads : 256) / 64);
        const int want = (slot_items_at(big.H, big.S, big.k, W == 1 ? 1 : 2) + waves_per - 1) / waves_per;
        const int most = W == 1 ? 128 : 48;
        int T = cfg.team_blocks > 0 ? cfg.team_blocks : (W >= cus ? 1 : std::max(1, std::min({resident / W, most, want})));
        if (cfg.blocks > 0 && cfg.team_blocks <= 0)
            T = std::min(T, cfg.blocks); // (a launch of fewer workgroups than the automatic team)
        int teams = cfg.blocks > 0 ? cfg.blocks / T : std::min(W, resident / T);
        if (teams < 1)
            return fail(TW_EINVAL, "%d teams of %d workgroups do not fit (%d resident workgroups)", teams, T, resident);
        team_T.assign(teams, T);
        team_q.assign(teams, 0);
    }
    // the queues' worlds, in order
    std::vector<int32_t> worder, wq_off(nq), wq_len(nq, 0), q_teams(nq, 0);
    for (int q = 0; q < nq; q++)
    {
        wq_off[q] = (int32_t)worder.size();
        for (int i : order)
            if (world_q[i] == q)
                worder.push_back(i);
        wq_len[q] = (int32_t)worder.size() - wq_off[q];
    }
    const int nteams = (int)team_T.size();
    for (int j = 0; j < nteams; j++)
        q_teams[team_q[j]]++;
    for (int q = 0; q < nq; q++)
        if (wq_len[q] && !q_teams[q])
            return fail(TW_EINVAL, "plan: queue %d has %d worlds and no team", q, wq_len[q]);
    std::vector<TeamPlan> plan(nteams);
    std::vector<int32_t> team_of;
    int blocks = 0;
    sh.teams = 0;
    for (int j = 0; j < nteams; j++)
    {
        TeamPlan &tp = plan[j];
        tp.T = team_T[j];
        tp.queue = team_q[j];
        if (tp.T > 255)
            return fail(TW_EINVAL, "teams of at most 255 workgroups");
        // hosts per wave item: as asked, else chosen per world by the kernel (about one work item
        // per wave of the team)
        tp.hpw = cfg.hosts_per_wave > 0 ? std::min(64, cfg.hosts_per_wave) : 0;
        blocks += tp.T;
        sh.teams |= tp.T > 1;
    }
    place_teams(plan, team_of);
    if (sh.teams && blocks > resident)
        return fail(TW_EINVAL, "the plan's %d workgroups exceed the %d the GPU holds at once (a team's workgroups must all be resident)",
                    blocks, resident);
    sh.blocks = blocks;
    sh.slot_budget = cfg.slot_budget > 0 ? cfg.slot_budget : 200000;
    UP(world_order, worder);
    UP(wq_off, wq_off);
    UP(wq_len, wq_len);
    ZERO(queue_next, nq);
    UP(team_of, team_of);
    UP(plan, plan);
    B.nwq = nq;
    ZERO(teams, nteams);
    const size_t waves = (size_t)sh.blocks * (sh.threads / 64);
    B.scr_waves = (int32_t)waves;
    ZERO(scr_ts, waves * 64 * 2 * MAX_TX);
    ZERO(scr_ref, waves * 64 * 2 * MAX_TX);
    ZERO(scr_up, waves * 64 * 2 * MAX_TX);
    b.plan = plan;
    b.nwq = nq;
    b.teams = B.teams;
    b.nteams = nteams;
#undef UP
#undef ZERO
    b.queue_next = B.queue_next;
    b.run = B.run;
    b.mg_est = B.mg_est;
    b.mg_done = B.mg_done;
    b.run_host = run;
    b.slots_before.assign(run.size(), 0);
    // the record as uploaded, and where its scratch pointers sit (tw_batch_run_noalloc points
    // them into the library's static scratch)
    static_assert(sizeof(Batch) <= NOALLOC_RECORD, "the batch record outgrew its place in the static scratch");
    b.batch_host.assign((const char *)&B, (const char *)&B + sizeof B);
    b.off_teams = offsetof(Batch, teams);
    b.off_queue_next = offsetof(Batch, queue_next);
    b.off_scr_ts = offsetof(Batch, scr_ts);
    b.off_scr_ref = offsetof(Batch, scr_ref);
    b.off_scr_up = offsetof(Batch, scr_up);
    Batch *bd = nullptr;
    if ((rc = dev_alloc(b, &bd, 1)))
        return rc;
    HIPCHK(hipMemcpy(bd, &B, sizeof B, hipMemcpyHostToDevice));
    b.batch_dev = bd;
    return TW_OK;
}

extern "C" int tw_batch_create(tw_world *const *worlds, int n, const tw_pack_cfg *cfg, tw_batch **out)
{
    if (!worlds || n <= 0 || !out)
        return fail(TW_EINVAL, "no worlds");
    tw_pack_cfg c;
    memset(&c, 0, sizeof c);
    if (cfg)
        c = *cfg;
    const int runs = c.chain_runs ? c.chain_runs : 9;
    if (runs != 9 && runs != 25)
        return fail(TW_EINVAL, "chain_runs: 9 or 25");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TW_EDEVICE, "no GPU: libtwins_core runs the model on the GPU only");
    std::unique_ptr<tw_batch> b(new tw_batch);
    b->runs = runs;
    b->no_graph = c.no_graph != 0;
    for (int i = 0; i < n; i++)
    {
        if (!worlds[i])
            return fail(TW_EINVAL, "world %d is null", i);
        b->worlds.emplace_back(new tw_world(*worlds[i]));
    }
    int rc = runs == 25 ? pack<25>(*b, c) : pack<9>(*b, c);
    if (rc)
        return rc;
    HIPCHK(hipStreamCreate(&b->stream));
    HIPCHK(hipEventCreate(&b->ev0));
    HIPCHK(hipEventCreate(&b->ev1));
    *out = b.release();
    return TW_OK;
}

// the launch (queue and team reset, then the persistent kernel) captured once as a HIP graph and
// replayed for every launch; a failed capture leaves the batch launching directly
static void capture(tw_batch &b)
{
    b.graph_failed = true;
    if (hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    {
        (void)hipGetLastError();
        return;
    }
    hipError_t e1 = tcp_reset((uint32_t *)b.queue_next, b.nwq, b.teams, sizeof(TeamState) * b.nteams, b.stream);
    hipError_t e2 = e1 == hipSuccess ? tcp_launch(b.runs, b.batch_dev, b.shape, b.stream) : e1;
    hipGraph_t g = nullptr;
    hipError_t e3 = hipStreamEndCapture(b.stream, &g);
    if (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess && g)
    {
        hipGraphExec_t x = nullptr;
        if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess && x)
        {
            b.graph = g;
            b.gexec = x;
            b.graph_failed = false;
            return;
        }
    }
    if (g)
        (void)hipGraphDestroy(g);
    (void)hipGetLastError();
}

extern "C" int tw_batch_run(tw_batch *b, tw_run_stats *stats)
{
    if (!b)
        return fail(TW_EINVAL, "null batch");
    const int W = (int)b->worlds.size();
    tw_run_stats st;
    memset(&st, 0, sizeof st);
    HIPCHK(hipDeviceSynchronize());
    std::vector<int64_t> slots_before(W);
    for (int i = 0; i < W; i++)
        slots_before[i] = b->run_host[i].slots_run;
    auto t0 = std::chrono::steady_clock::now();
    for (;;)
    {
        int unfinished = 0;
        for (auto &r : b->run_host)
            unfinished += !r.finished;
        if (!unfinished)
            break;
        if (!b->gexec && !b->no_graph && !b->graph_failed)
            capture(*b);
        HIPCHK(hipEventRecord(b->ev0, b->stream));
        if (b->gexec)
            HIPCHK(hipGraphLaunch(b->gexec, b->stream));
        else
        {
            HIPCHK(tcp_reset((uint32_t *)b->queue_next, b->nwq, b->teams, sizeof(TeamState) * b->nteams, b->stream));
            HIPCHK(tcp_launch(b->runs, b->batch_dev, b->shape, b->stream));
        }
        HIPCHK(hipEventRecord(b->ev1, b->stream));
        HIPCHK(hipMemcpyAsync(b->run_host.data(), b->run, sizeof(WorldRun) * W, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        st.kernel_s += ms / 1000.0;
        st.launches++;
        // every launch runs at least one slot of every unfinished world it reaches
        int64_t ran = 0;
        for (int i = 0; i < W; i++)
            ran += b->run_host[i].slots_run - slots_before[i];
        if (ran == 0)
            return fail(TW_ESTATE, "a launch ran no slot of the %d unfinished worlds", unfinished);
        for (int i = 0; i < W; i++)
            slots_before[i] = b->run_host[i].slots_run;
    }
    st.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto &r : b->run_host)
        if (r.error)
            return fail(TW_EDEVICE, "a team barrier timed out: the world's results are incomplete");
    for (auto &r : b->run_host)
    {
        st.slots_max = std::max<int64_t>(st.slots_max, r.slots_run);
        st.packet_events += r.ev[EV_PKT_EVENTS];
        st.flows_done += r.flows_done;
        st.worlds_done += r.finished != 0;
    }
    st.graph = b->gexec != nullptr;
    if (stats)
        *stats = st;
    return TW_OK;
}

// tw_batch_run with the launch's scratch memory (the batch record the kernel reads, the teams'
// barrier words, the world-queue counters and the switch groups' sort space) in the library's
// static device array instead of the batch's own: no allocation of any kind, no graph. Scratch
// layout: the record, then each region on a 256-byte boundary.
extern "C" int tw_batch_run_noalloc(tw_batch *b, tw_run_stats *stats)
{
    if (!b)
        return fail(TW_EINVAL, "null batch");
    if (b->for_server || b->batch_host.empty())
        return fail(TW_EINVAL, "tw_batch_run_noalloc: a batch made by tw_batch_create");
    void *sp = nullptr;
    HIPCHK(tcp_static_scratch(&sp));
    const size_t per_wave = (size_t)64 * 2 * MAX_TX, waves = (size_t)b->shape.blocks * (b->shape.threads / 64);
    size_t end = 0;
    auto region = [&](size_t bytes) {
        const size_t at = (end + 255) & ~(size_t)255;
        end = at + bytes;
        return at;
    };
    const size_t o_rec = region(NOALLOC_RECORD), o_teams = region(sizeof(TeamState) * std::max(1, b->nteams)),
                 o_queue = region(4 * (size_t)std::max(1, b->nwq)), o_ts = region(8 * waves * per_wave), o_ref = region(4 * waves * per_wave),
                 o_up = region(waves * per_wave);
    if (end > TW_SCRATCH_BYTES)
        return fail(TW_EINVAL,
                    "tw_batch_run_noalloc: the batch's launch (%d workgroups of %d threads) needs %zu bytes of scratch, the static "
                    "scratch holds %zu (at most %zu wavefronts): make the batch with fewer workgroups (tw_pack_cfg.blocks)",
                    b->shape.blocks, b->shape.threads, end, TW_SCRATCH_BYTES, (size_t)(TW_SCRATCH_BYTES - 2 * NOALLOC_RECORD) / (13 * per_wave));
    char *base = (char *)sp;
    // the record with its scratch pointers moved into the static array (a stack copy: no allocation)
    alignas(64) char rec[NOALLOC_RECORD];
    memcpy(rec, b->batch_host.data(), b->batch_host.size());
    auto put = [&](size_t field, size_t at) {
        void *v = base + at;
        memcpy(rec + field, &v, sizeof v);
    };
    put(b->off_teams, o_teams);
    put(b->off_queue_next, o_queue);
    put(b->off_scr_ts, o_ts);
    put(b->off_scr_ref, o_ref);
    put(b->off_scr_up, o_up);
    HIPCHK(hipMemcpyAsync(base + o_rec, rec, b->batch_host.size(), hipMemcpyHostToDevice, b->stream));
    const int W = (int)b->worlds.size();
    tw_run_stats st;
    memset(&st, 0, sizeof st);
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < W; i++)
        b->slots_before[i] = b->run_host[i].slots_run;
    auto t0 = std::chrono::steady_clock::now();
    for (;;)
    {
        int unfinished = 0;
        for (auto &r : b->run_host)
            unfinished += !r.finished;
        if (!unfinished)
            break;
        HIPCHK(hipEventRecord(b->ev0, b->stream));
        // counters and barrier words start at zero every launch (a fill kernel, no memset)
        HIPCHK(tcp_reset((uint32_t *)(base + o_queue), b->nwq, base + o_teams, sizeof(TeamState) * b->nteams, b->stream));
        HIPCHK(tcp_launch(b->runs, base + o_rec, b->shape, b->stream));
        HIPCHK(hipEventRecord(b->ev1, b->stream));
        HIPCHK(hipMemcpyAsync(b->run_host.data(), b->run, sizeof(WorldRun) * W, hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        st.kernel_s += ms / 1000.0;
        st.launches++;
        int64_t ran = 0;
        for (int i = 0; i < W; i++)
        {
            ran += b->run_host[i].slots_run - b->slots_before[i];
            b->slots_before[i] = b->run_host[i].slots_run;
        }
        if (ran == 0)
            return fail(TW_ESTATE, "a launch ran no slot of the %d unfinished worlds", unfinished);
    }
    st.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto &r : b->run_host)
        if (r.error)
            return fail(TW_EDEVICE, "a team barrier timed out: the world's results are incomplete");
    for (auto &r : b->run_host)
    {
        st.slots_max = std::max<int64_t>(st.slots_max, r.slots_run);
        st.packet_events += r.ev[EV_PKT_EVENTS];
        st.flows_done += r.flows_done;
        st.worlds_done += r.finished != 0;
    }
    if (stats)
        *stats = st;
    return TW_OK;
}

extern "C" int tw_fct_get(tw_batch *b, int world, int64_t *est_ns, int64_t *done_ns)
{
    if (!b || world < 0 || world >= (int)b->worlds.size())
        return fail(TW_EINVAL, "no world %d", world);
    const tw_world &w = *b->worlds[world];
    std::vector<int64_t> est(w.M), done(w.M);
    if (w.M)
    {
        HIPCHK(hipMemcpy(est.data(), b->mg_est + b->msg0[world], 8 * w.M, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(done.data(), b->mg_done + b->msg0[world], 8 * w.M, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < w.M; i++)
    {
        int j = w.mnew[i];
        if (est_ns)
            est_ns[i] = est[j] >= 0 ? est[j] / 1000 : -1;
        if (done_ns)
            done_ns[i] = done[j] >= 0 ? done[j] / 1000 : -1;
    }
    return TW_OK;
}

extern "C" int tw_world_stats_get(tw_batch *b, int world, tw_world_stats *st)
{
    if (!b || !st || world < 0 || world >= (int)b->worlds.size())
        return fail(TW_EINVAL, "no world %d", world);
    const WorldRun &r = b->run_host[world];
    st->slots = r.slots_run;
    st->device_s = (double)r.ticks / 100e6; // the 100 MHz constant clock
    st->team_blocks = r.team;
    st->finished = r.finished;
    st->packet_events = r.ev[EV_PKT_EVENTS];
    return TW_OK;
}

extern "C" int tw_launch_capacity(int chain_runs, int block_threads, int32_t *resident_blocks)
{
    if (!resident_blocks)
        return fail(TW_EINVAL, "null output");
    const int threads = block_threads > 0 ? block_threads : 256;
    if (threads % 64 || threads > 256)
        return fail(TW_EINVAL, "block_threads: a multiple of 64 up to 256");
    int ndev = 0, dev = 0, cus = 0, per_cu = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TW_EDEVICE, "no GPU: libtwins_core runs the model on the GPU only");
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(tcp_max_blocks(chain_runs == 25 ? 25 : 9, threads, &per_cu));
    *resident_blocks = std::max(1, per_cu) * cus;
    return TW_OK;
}

extern "C" int tw_counters_get(tw_batch *b, int world, uint64_t ev[32])
{
    if (!b || !ev || world < 0 || world >= (int)b->worlds.size())
        return fail(TW_EINVAL, "no world %d", world);
    const WorldRun &r = b->run_host[world];
    for (int i = 0; i < 32; i++)
        ev[i] = r.ev[i];
    ev[31] = (uint64_t)r.slot;
    return TW_OK;
}

// ------------------------------------------------------------------ the queue trace
extern "C" int tw_qtrace_enable(tw_batch *b, int64_t from_ps, int64_t until_ps, uint64_t max_records)
{
    static_assert(offsetof(BatchT<9>, qt) == offsetof(BatchT<25>, qt), "one trace offset");
    if (!b || b->for_server)
        return fail(TW_EINVAL, "qtrace: a batch of its own");
    if (b->qt.rec)
        return fail(TW_EINVAL, "qtrace: already enabled");
    QTrace q{};
    q.cap = max_records ? max_records : 20000000;
    q.from = from_ps;
    q.until = until_ps;
    int rc;
    if ((rc = dev_alloc(*b, &q.rec, q.cap)) || (rc = dev_zero(*b, &q.cnt, 1)))
        return rc;
    HIPCHK(hipMemcpy((char *)b->batch_dev + offsetof(BatchT<9>, qt), &q, sizeof q, hipMemcpyHostToDevice));
    b->qt = q;
    return TW_OK;
}

extern "C" int tw_qtrace_write(tw_batch *b, const char *path, uint64_t *records)
{
    if (!b || !path)
        return fail(TW_EINVAL, "qtrace: no batch or path");
    if (!b->qt.rec)
        return fail(TW_EINVAL, "qtrace: not enabled (tw_qtrace_enable)");
    unsigned long long n = 0;
    HIPCHK(hipMemcpy(&n, b->qt.cnt, sizeof n, hipMemcpyDeviceToHost));
    n = std::min<unsigned long long>(n, b->qt.cap);
    std::vector<QRec> q(n);
    if (n)
        HIPCHK(hipMemcpy(q.data(), b->qt.rec, n * sizeof(QRec), hipMemcpyDeviceToHost));
    // world-local numbering: node and device of the reference's queue id (hosts 0..H-1 on device 1,
    // then switch x's port p as node H + x on device p + 1), hosts of the connection
    struct Line
    {
        int64_t ts;
        int32_t world, node, dev, src, dst;
        uint32_t sa;
        int qlen;
        const char *kind;
    };
    std::vector<Line> out(n);
    std::vector<int32_t> link0(b->worlds.size() + 1, 0);
    for (size_t i = 0; i < b->worlds.size(); i++)
        link0[i + 1] = link0[i] + b->worlds[i]->L;
    for (size_t i = 0; i < n; i++)
    {
        const QRec &r = q[i];
        const int wi = (int)(std::upper_bound(link0.begin(), link0.end(), r.link) - link0.begin()) - 1;
        if (wi < 0 || wi >= (int)b->worlds.size())
            return fail(TW_EDEVICE, "qtrace: record %zu names link %d", i, r.link);
        const tw_world &w = *b->worlds[wi];
        const int l = r.link - link0[wi];
        const uint32_t c = r.flow - (uint32_t)b->conn0[wi];
        if (c >= (uint32_t)w.NC)
            return fail(TW_EDEVICE, "qtrace: record %zu names connection %u", i, r.flow);
        Line &o = out[i];
        o.ts = r.ts;
        o.world = wi;
        o.node = l < w.H ? l : w.H + (l - w.H) / w.k;
        o.dev = l < w.H ? 1 : (l - w.H) % w.k + 1;
        o.src = r.to_sender ? w.cs[c].dst : w.cs[c].src;
        o.dst = r.to_sender ? w.cs[c].src : w.cs[c].dst;
        o.sa = r.sa;
        o.qlen = r.qlen;
        o.kind = ((r.flags & F_SYN) && (r.flags & F_ACK)) ? "SA" : (r.flags & F_DATA) ? "D" : (r.flags & F_SYN) ? "S" : "A";
    }
    std::stable_sort(out.begin(), out.end(), [](const Line &a, const Line &b) {
        if (a.ts != b.ts)
            return a.ts < b.ts;
        if (a.node != b.node)
            return a.node < b.node;
        if (a.dev != b.dev)
            return a.dev < b.dev;
        return a.world < b.world;
    });
    FILE *f = fopen(path, "w");
    if (!f)
        return fail(TW_EIO, "cannot write %s", path);
    for (const Line &o : out)
        fprintf(f, "%lld %d %d %d %d %s %u %d\n", (long long)o.ts, o.node, o.dev, o.src, o.dst, o.kind, o.sa, o.qlen);
    if (fclose(f))
        return fail(TW_EIO, "cannot write %s", path);
    if (records)
        *records = n;
    return TW_OK;
}

extern "C" int tw_batch_destroy(tw_batch *b)
{
    delete b;
    return TW_OK;
}

// ------------------------------------------------------------------ the serving loop
// Requests (one world each) are admitted in arrival order while their device footprint fits the
// arena, packed there, and published to the persistent server kernel's queue for their size; the
// kernel runs them side by side on its teams and flags each done in pinned host memory. The host
// side (tw_server_poll) collects completions, admits waiting requests as memory frees up, and
// relaunches the kernel when it has left (idle) while requests wait. No counterpart in the
// reference, whose runs are one process per simulation.
namespace
{
struct SrvReq
{
    std::unique_ptr<tw_world> w;  // until admitted
    std::unique_ptr<tw_batch> b;  // once admitted, until its result is taken
    int queue = -1;
    uint32_t index = 0;
    int state = TW_REQ_WAITING;
};
} // namespace

struct tw_server
{
    int runs = 9, threads = 256, blocks = 0, nq = 3, cap = 0;
    Arena arena;
    hipStream_t ks = nullptr, up = nullptr;
    ServerCtl *ctl_d = nullptr;
    ServerReq *req = nullptr; // pinned
    unsigned *tail = nullptr;
    int *shutdown = nullptr, *done = nullptr;
    unsigned *next_d = nullptr;
    TeamState *teams_d = nullptr;
    int nteams = 0;
    int64_t small_max = 512, mid_max = 4096; // queue by entities: small, mid, large
    int q_teams[3] = {0, 0, 0};              // teams per queue
    std::vector<unsigned> published;
    std::vector<SrvReq> reqs;
    std::vector<int> waiting; // ids, in arrival order
    size_t wait_head = 0;
    std::vector<void *> dev, host;
    std::vector<std::unique_ptr<tw_batch>> graveyard; // failed requests' batches, freed once the kernel has left
    bool launched = false;
    uint64_t relaunches = 0;
    ~tw_server()
    {
        if (shutdown)
            __atomic_store_n(shutdown, 1, __ATOMIC_RELEASE);
        if (ks)
            (void)hipStreamSynchronize(ks);
        graveyard.clear();
        reqs.clear();
        for (void *p : dev)
            (void)hipFree(p);
        for (void *p : host)
            (void)hipHostFree(p);
        if (ks)
            (void)hipStreamDestroy(ks);
        if (up)
            (void)hipStreamDestroy(up);
    }
};

template <typename T>
static int srv_dev(tw_server &s, T **out, size_t n)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess)
        return fail(TW_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
    s.dev.push_back(p);
    *out = (T *)p;
    return TW_OK;
}
template <typename T>
static int srv_pinned(tw_server &s, T **out, size_t n)
{
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, sizeof(T) * std::max<size_t>(n, 1), hipHostMallocMapped);
    if (e != hipSuccess)
        return fail(TW_ENOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
    memset(p, 0, sizeof(T) * std::max<size_t>(n, 1));
    s.host.push_back(p);
    *out = (T *)p;
    return TW_OK;
}
template <typename T>
static T *mapped(T *p)
{
    void *d = nullptr;
    return hipHostGetDevicePointer(&d, p, 0) == hipSuccess ? (T *)d : p;
}

extern "C" int tw_server_create(const tw_server_cfg *cfg, tw_server **out)
{
    if (!out)
        return fail(TW_EINVAL, "null output");
    tw_server_cfg c;
    memset(&c, 0, sizeof c);
    if (cfg)
        c = *cfg;
    int ndev = 0, dev = 0, cus = 0, per_cu = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TW_EDEVICE, "no GPU: libtwins_core runs the model on the GPU only");
    {
        // HIP maps the process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless set);
        // a stream sharing a queue with this server's persistent kernel waits until it leaves
        const char *q = getenv("GPU_MAX_HW_QUEUES");
        static bool told = false;
        if ((!q || atoi(q) < 8) && !told)
        {
            told = true;
            fprintf(stderr, "twins: serving loop with %s hardware queues for HIP streams: work on a stream that shares one with "
                            "the server's persistent kernel waits for it to leave; set GPU_MAX_HW_QUEUES=16 before the first HIP call\n",
                    q ? q : "4 (the default)");
        }
    }
    std::unique_ptr<tw_server> s(new tw_server);
    s->runs = c.chain_runs == 25 ? 25 : 9;
    s->threads = c.block_threads > 0 ? c.block_threads : 256;
    if (s->threads % 64 || s->threads > 256)
        return fail(TW_EINVAL, "block_threads: a multiple of 64 up to 256");
    s->cap = c.max_requests > 0 ? c.max_requests : 65536;
    // an idle team gives its claim back after idle_s; a claim is held at most SERVER_WAIT_S in all
    if (c.idle_s >= SERVER_WAIT_S)
        return fail(TW_EINVAL, "idle_s %.1f: less than %.0f s", c.idle_s, SERVER_WAIT_S);
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(tcp_max_blocks(s->runs, s->threads, &per_cu));
    const int resident = std::max(1, per_cu) * cus;
    // teams: large worlds on teams of 16 workgroups, mid-size on 4, small on one; a share of the
    // GPU stays free for the uploads' fill kernels
    const int nl = c.large_teams > 0 ? c.large_teams : 4, nm = c.mid_teams > 0 ? c.mid_teams : 8;
    const int headroom = cus / 8;
    const int zm = c.mid_size > 0 ? c.mid_size : 4, zl = c.large_size > 0 ? c.large_size : 16;
    if (zm > 255 || zl > 255)
        return fail(TW_EINVAL, "server teams of at most 255 workgroups");
    const int ns = c.small_teams > 0 ? c.small_teams : resident - headroom - zl * nl - zm * nm;
    if (ns < 1 || ns + zm * nm + zl * nl > resident)
        return fail(TW_EINVAL, "server teams (%d + %d x %d + %d x %d workgroups) do not fit the %d the GPU holds", ns, zm, nm, zl, nl,
                    resident);
    std::vector<TeamPlan> plan;
    std::vector<int32_t> team_of;
    auto add = [&](int n, int T, int q) {
        for (int i = 0; i < n; i++)
            plan.push_back(TeamPlan{0, T, 0, q});
    };
    add(ns, 1, 0);
    add(nm, zm, 1);
    add(nl, zl, 2);
    s->q_teams[0] = ns, s->q_teams[1] = nm, s->q_teams[2] = nl;
    place_teams(plan, team_of);
    s->blocks = (int)team_of.size();
    s->nteams = (int)plan.size();
    // the arena
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    const size_t want = c.arena_bytes ? c.arena_bytes : fr / 10 * 6;
    void *ab = nullptr;
    if (hipMalloc(&ab, want) != hipSuccess)
        return fail(TW_ENOMEM, "arena of %zu bytes", want);
    s->dev.push_back(ab);
    s->arena.base = (char *)ab;
    s->arena.size = want;
    s->arena.free_[0] = want;
    HIPCHK(hipStreamCreateWithFlags(&s->ks, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking));
    int rc;
    if ((rc = srv_pinned(*s, &s->req, (size_t)s->nq * s->cap)) || (rc = srv_pinned(*s, &s->tail, s->nq)) ||
        (rc = srv_pinned(*s, &s->shutdown, 1)) || (rc = srv_pinned(*s, &s->done, s->cap)))
        return rc;
    ServerCtl h;
    memset(&h, 0, sizeof h);
    h.req = mapped(s->req);
    h.tail = mapped(s->tail);
    h.shutdown = mapped(s->shutdown);
    h.done = mapped(s->done);
    h.nq = s->nq;
    h.cap = s->cap;
    h.idle_ticks = (unsigned long long)((c.idle_s > 0 ? c.idle_s : 2.0) * 100e6);
    const size_t waves = (size_t)s->blocks * (s->threads / 64);
    if ((rc = srv_dev(*s, &s->next_d, s->nq)) || (rc = srv_dev(*s, &s->teams_d, s->nteams)) || (rc = srv_dev(*s, &h.team_of, team_of.size())) ||
        (rc = srv_dev(*s, &h.plan, plan.size())) || (rc = srv_dev(*s, &h.scr_ts, waves * 64 * 2 * MAX_TX)) ||
        (rc = srv_dev(*s, &h.scr_ref, waves * 64 * 2 * MAX_TX)) || (rc = srv_dev(*s, &h.scr_up, waves * 64 * 2 * MAX_TX)) ||
        (rc = srv_dev(*s, &s->ctl_d, 1)))
        return rc;
    h.next = s->next_d;
    h.teams = s->teams_d;
    HIPCHK(hipMemcpy(h.team_of, team_of.data(), 4 * team_of.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h.plan, plan.data(), sizeof(TeamPlan) * plan.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->next_d, 0, 4 * s->nq));
    HIPCHK(hipMemset(s->teams_d, 0, sizeof(TeamState) * s->nteams));
    HIPCHK(hipMemcpy(s->ctl_d, &h, sizeof h, hipMemcpyHostToDevice));
    s->published.assign(s->nq, 0);
    *out = s.release();
    return TW_OK;
}

// resident requests a serving queue holds per team of it
constexpr int SRV_DEPTH = 4;

// completions, admissions, and a relaunch of the kernel when it has left while requests wait
static int srv_service(tw_server &s)
{
    for (auto &r : s.reqs)
        if (r.state == TW_REQ_RESIDENT)
        {
            int d = __atomic_load_n(&s.done[&r - &s.reqs[0]], __ATOMIC_ACQUIRE);
            if (d)
                r.state = d == 1 ? TW_REQ_DONE : TW_REQ_FAILED;
        }
    // Admissions, in arrival order within each queue. A queue holds at most SRV_DEPTH resident
    // requests per team (more would only wait for its teams while holding device memory others
    // could run in), and a request that is over its queue's share or does not fit the arena is
    // passed over by later ones of other queues, so a large world waiting for memory or for the
    // large teams does not hold up small ones.
    int resident_q[3] = {0, 0, 0};
    for (auto &r : s.reqs)
        if (r.state == TW_REQ_RESIDENT)
            resident_q[r.queue]++;
    bool blocked[3] = {false, false, false};
    for (size_t i = s.wait_head; i < s.waiting.size(); i++)
    {
        const int id = s.waiting[i];
        SrvReq &r = s.reqs[id];
        if (r.state != TW_REQ_WAITING)
            continue;
        const int64_t n = (int64_t)r.w->H + r.w->S;
        const int q = n <= s.small_max ? 0 : n <= s.mid_max ? 1 : 2;
        if (blocked[q])
            continue; // (in order within a queue)
        if (resident_q[q] >= SRV_DEPTH * std::max(1, s.q_teams[q]) ||
            world_bytes(*r.w, s.runs) + (64u << 20) > s.arena.size - s.arena.used)
        {
            blocked[q] = true;
            continue;
        }
        std::unique_ptr<tw_batch> b(new tw_batch);
        b->runs = s.runs;
        b->arena = &s.arena;
        b->up = s.up;
        b->for_server = true;
        b->worlds.emplace_back(new tw_world(*r.w));
        tw_pack_cfg pc;
        memset(&pc, 0, sizeof pc);
        int rc = s.runs == 25 ? pack<25>(*b, pc) : pack<9>(*b, pc);
        if (rc == TW_ENOMEM)
        {
            blocked[q] = true; // the arena is fragmented: wait for more to free
            continue;
        }
        if (rc)
            return rc;
        resident_q[q]++;
        r.queue = q;
        r.index = s.published[r.queue];
        if ((int)r.index >= s.cap)
            return fail(TW_ESTATE, "queue %d is full (%d requests)", r.queue, s.cap);
        s.req[(size_t)r.queue * s.cap + r.index] = ServerReq{b->batch_dev, id, 0};
        r.b = std::move(b);
        r.w.reset();
        r.state = TW_REQ_RESIDENT;
        s.published[r.queue] = r.index + 1;
        __atomic_store_n(&s.tail[r.queue], r.index + 1, __ATOMIC_RELEASE);
    }
    while (s.wait_head < s.waiting.size() && s.reqs[s.waiting[s.wait_head]].state != TW_REQ_WAITING)
        s.wait_head++;
    bool unserved = false;
    for (auto &r : s.reqs)
        unserved |= r.state == TW_REQ_RESIDENT;
    if (!unserved)
        return TW_OK;
    if (s.launched)
    {
        hipError_t q = hipStreamQuery(s.ks);
        if (q == hipErrorNotReady)
            return TW_OK;
        HIPCHK(q);
        // the kernel has left: look again at the flags it set before restarting the counters
        for (auto &r : s.reqs)
            if (r.state == TW_REQ_RESIDENT)
            {
                int d = __atomic_load_n(&s.done[&r - &s.reqs[0]], __ATOMIC_ACQUIRE);
                if (d)
                    r.state = d == 1 ? TW_REQ_DONE : TW_REQ_FAILED;
            }
        s.relaunches++;
    }
    // no kernel runs now: failed requests' memory may go back to the arena
    s.graveyard.clear();
    // each queue restarts after its last published index, with the requests not yet run published
    // again there (a request the old kernel finished is never claimed twice, and its memory may
    // already be back in the arena)
    std::vector<unsigned> first(s.nq, 0); // (a first launch: every published request is still to run)
    if (s.launched)
    {
        first = s.published;
        for (auto &r : s.reqs)
            if (r.state == TW_REQ_RESIDENT)
            {
                const unsigned i = s.published[r.queue];
                if ((int)i >= s.cap)
                    return fail(TW_ESTATE, "queue %d is full (%d requests)", r.queue, s.cap);
                s.req[(size_t)r.queue * s.cap + i] = s.req[(size_t)r.queue * s.cap + r.index];
                r.index = i;
                s.published[r.queue] = i + 1;
            }
    }
    for (int q = 0; q < s.nq; q++)
        __atomic_store_n(&s.tail[q], s.published[q], __ATOMIC_RELEASE);
    std::vector<TeamState> zero(s.nteams);
    memset(zero.data(), 0, sizeof(TeamState) * s.nteams);
    HIPCHK(hipMemcpyAsync(s.next_d, first.data(), 4 * s.nq, hipMemcpyHostToDevice, s.up));
    HIPCHK(hipMemcpyAsync(s.teams_d, zero.data(), sizeof(TeamState) * s.nteams, hipMemcpyHostToDevice, s.up));
    HIPCHK(hipStreamSynchronize(s.up));
    HIPCHK(tcp_server_launch(s.runs, s.ctl_d, s.blocks, s.threads, s.ks));
    s.launched = true;
    return TW_OK;
}

extern "C" int tw_server_submit(tw_server *s, const tw_world *w, int32_t *id)
{
    if (!s || !w || !id)
        return fail(TW_EINVAL, "null argument");
    if ((int)s->reqs.size() >= s->cap)
        return fail(TW_ESTATE, "the server took its %d requests", s->cap);
    if (world_bytes(*w, s->runs) + (64u << 20) > s->arena.size)
        return fail(TW_ENOMEM, "the world needs more device memory than the arena holds");
    SrvReq r;
    r.w.reset(new tw_world(*w));
    s->reqs.push_back(std::move(r));
    *id = (int32_t)s->reqs.size() - 1;
    s->waiting.push_back(*id);
    return srv_service(*s);
}

extern "C" int tw_server_poll(tw_server *s, tw_server_stats *st)
{
    if (!s)
        return fail(TW_EINVAL, "null server");
    int rc = srv_service(*s);
    if (rc)
        return rc;
    if (st)
    {
        memset(st, 0, sizeof *st);
        for (auto &r : s->reqs)
        {
            st->waiting += r.state == TW_REQ_WAITING;
            st->resident += r.state == TW_REQ_RESIDENT;
            st->done += r.state == TW_REQ_DONE || r.state == TW_REQ_TAKEN;
            st->failed += r.state == TW_REQ_FAILED;
        }
        st->arena_used = s->arena.used;
        st->arena_bytes = s->arena.size;
        st->relaunches = s->relaunches;
    }
    return TW_OK;
}

extern "C" int tw_server_state(tw_server *s, int32_t id, int32_t *state)
{
    if (!s || !state || id < 0 || id >= (int)s->reqs.size())
        return fail(TW_EINVAL, "no request %d", id);
    *state = s->reqs[id].state;
    return TW_OK;
}

extern "C" int tw_server_result(tw_server *s, int32_t id, int64_t *est_ns, int64_t *done_ns, uint64_t ev[32])
{
    if (!s || id < 0 || id >= (int)s->reqs.size())
        return fail(TW_EINVAL, "no request %d", id);
    SrvReq &r = s->reqs[id];
    if (r.state != TW_REQ_DONE && r.state != TW_REQ_FAILED)
        return fail(TW_ESTATE, "request %d is not finished (state %d)", id, r.state);
    tw_batch &b = *r.b;
    HIPCHK(hipMemcpyAsync(b.run_host.data(), b.run, sizeof(WorldRun), hipMemcpyDeviceToHost, s->up));
    HIPCHK(hipStreamSynchronize(s->up));
    int rc = tw_fct_get(&b, 0, est_ns, done_ns);
    if (rc)
        return rc;
    if (ev)
        for (int i = 0; i < 32; i++)
            ev[i] = b.run_host[0].ev[i];
    if (ev)
        ev[31] = b.run_host[0].slot;
    const bool failed = r.state == TW_REQ_FAILED;
    if (failed && s->launched && hipStreamQuery(s->ks) == hipErrorNotReady)
        s->graveyard.push_back(std::move(r.b)); // (the team's other workgroups may still write into it)
    r.b.reset(); // its memory goes back to the arena
    r.state = TW_REQ_TAKEN;
    return failed ? fail(TW_EDEVICE, "request %d failed on the device (a team wait timed out)", id) : TW_OK;
}

extern "C" int tw_server_destroy(tw_server *s)
{
    delete s;
    return TW_OK;
}
