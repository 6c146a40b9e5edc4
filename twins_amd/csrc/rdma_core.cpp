// Host side of the RDMA model: HPCC config files, TwiNS topology and flow files, routes, the
// worlds' device state, the run loop and results. The reference does all of this in one main
// (src/rdma/rmain.cu); here each stage is a C-ABI call (include/twins_rdma.h).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "rdma_model.h"
#include "rdma_kernel.h"
#include "mrg_host.h"
#include "arena.h"
#include "../../include/twins.h"
#include "../../include/twins_rdma.h"

using std::string;
using std::vector;


static thread_local string g_err;
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
extern "C" const char *trd_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------ HPCC config (src/rdma/rmain.cu:100-161)
typedef std::map<string, vector<string>> Config;
static bool read_config(const string &path, Config &cfg)
{
    std::ifstream f(path);
    if (!f)
        return false;
    string line;
    while (std::getline(f, line))
    {
        std::istringstream is(line);
        string key, v;
        if (!(is >> key))
            continue;
        vector<string> vals;
        while (is >> v)
            vals.push_back(v);
        cfg[key] = vals;
    }
    return true;
}
static void set_override(Config &cfg, const string &k, const string &v)
{
    vector<string> vals;
    size_t b = 0;
    for (size_t c; (c = v.find(',', b)) != string::npos; b = c + 1)
        vals.push_back(v.substr(b, c - b));
    vals.push_back(v.substr(b));
    cfg[k] = vals;
}
struct Cfg
{
    const Config &c;
    bool has(const char *k) const { return c.count(k) && !c.at(k).empty(); }
    string str(const char *k, const string &d) const { return has(k) ? c.at(k)[0] : d; }
    double num(const char *k, double d) const { return has(k) ? atof(c.at(k)[0].c_str()) : d; }
    uint32_t u32(const char *k, uint32_t d) const { return has(k) ? (uint32_t)strtoul(c.at(k)[0].c_str(), 0, 10) : d; }
};
// ns-3 DataRate strings: a number and a unit (bps, Kbps, Mbps, Gbps or b/s forms)
static uint64_t parse_rate(const string &s)
{
    size_t i = 0;
    while (i < s.size() && (isdigit((unsigned char)s[i]) || s[i] == '.'))
        i++;
    double v = atof(s.substr(0, i).c_str());
    const char u = i < s.size() ? s[i] : 'b';
    double m = (u == 'K' || u == 'k') ? 1e3 : u == 'M' ? 1e6 : u == 'G' ? 1e9 : 1;
    return (uint64_t)(v * m);
}

// ------------------------------------------------------------------ a world on the host
struct trd_world
{
    rd::Params P;
    int H = 0, S = 0, K = 1;
    vector<int> deg, out_link, in_link; // [N K]
    vector<uint64_t> lbps;              // [L]
    vector<int64_t> ldelay;
    vector<rd::Route> rt;               // [S H]
    vector<uint8_t> rtp;
    vector<rd::PortCfg> pcfg;           // switch ports in turn
    vector<rd::FlowInfo> fi;
    vector<int> start_list, start_off;  // flows by (src, start order)
    uint64_t max_rtt = 0;
    uint32_t max_bdp = 0;
    double stop_ns = 0;
    int L() const { return (int)lbps.size(); }
};

// cpush for the host's short chains (no tails)
static void host_push(rd::Chain *c, int64_t lead, int ord)
{
    const rd::Chain par = *c;
    const uint32_t code = rd::ccode(lead, ord);
    if (par.r[0].code == code)
        c->r[0].cnt++;
    else
    {
        for (int i = rd::RUNS - 1; i > 0; i--)
            c->r[i] = par.r[i - 1];
        c->r[0].code = code;
        c->r[0].cnt = 1;
    }
}

extern "C" int trd_world_create(const char *config_path, const char *const *overrides, int32_t n_overrides, const char *topo_path,
                                const char *flow_path, uint64_t rate_bps, int64_t delay_ns, trd_world **out)
{
    if (!config_path || !topo_path || !flow_path || !out)
        return fail(TW_EINVAL, "trd_world_create: null argument");
    Config config;
    if (!read_config(config_path, config))
        return fail(TW_EIO, "cannot open %s", config_path);
    for (int i = 0; i < n_overrides; i++)
    {
        string kv = overrides[i];
        size_t eq = kv.find('=');
        if (eq == string::npos)
            return fail(TW_EINVAL, "bad override %s", kv.c_str());
        set_override(config, kv.substr(0, eq), kv.substr(eq + 1));
    }
    {
        const size_t n = strlen(flow_path);
        if (n > 4 && strcmp(flow_path + n - 4, ".dag") == 0)
            return fail(TW_EINVAL, "%s: DAG flow files are not modelled", flow_path);
    }
    const uint64_t bps = rate_bps ? rate_bps : 100000000000ull;
    const int64_t delay = delay_ns > 0 ? delay_ns : 1000;
    Cfg cfg{config};
    std::unique_ptr<trd_world> wp(new trd_world);
    trd_world &w = *wp;
    rd::Params &P = w.P;
    P = rd::Params{};
    // parameters and their defaults (src/rdma/rmain.cu:369-420)
    P.cc = (int)cfg.u32("CC_MODE", 1);
    if (P.cc != 1 && P.cc != 3 && P.cc != 8)
        return fail(TW_EINVAL, "CC_MODE %d: only 1 (DCQCN), 3 (HPCC) and 8 (DCTCP) are modelled", P.cc);
    if (cfg.u32("MULTI_RATE", 1) && P.cc == 3)
        return fail(TW_EINVAL, "MULTI_RATE 1 is not modelled");
    if (cfg.u32("L2_BACK_TO_ZERO", 0))
        return fail(TW_EINVAL, "L2_BACK_TO_ZERO 1 is not modelled");
    P.int_size = P.cc == 3 ? 42 : 0;
    P.mtu = cfg.u32("PACKET_PAYLOAD_SIZE", 1000);
    P.ack_interval = cfg.u32("L2_ACK_INTERVAL", 0);
    P.chunk = cfg.u32("L2_CHUNK_SIZE", 0);
    if (P.ack_interval == 0)
        return fail(TW_EINVAL, "L2_ACK_INTERVAL 0 is not modelled");
    P.nack_interval = 500000;
    P.var_win = (int)cfg.u32("VAR_WIN", 0);
    P.rate_bound = (int)cfg.u32("RATE_BOUND", 1);
    P.ack_high_prio = (int)cfg.u32("ACK_HIGH_PRIO", 0);
    P.ecn = (int)cfg.u32("ENABLE_QCN", 1);
    P.g = cfg.num("EWMA_GAIN", 1.0 / 16);
    P.rate_first_cnp = 1.0;
    P.clamp = (int)cfg.u32("CLAMP_TARGET_RATE", 0);
    // MicroSeconds of an integer: the attribute's double truncated first
    P.alpha_int = (int64_t)(uint64_t)cfg.num("ALPHA_RESUME_INTERVAL", 55) * 1000;
    P.dec_int = (int64_t)(uint64_t)cfg.num("RATE_DECREASE_INTERVAL", 4) * 1000;
    P.rp_int = (int64_t)(uint64_t)cfg.num("RP_TIMER", 1500) * 1000;
    P.rpg_threshold = cfg.u32("FAST_RECOVERY_TIMES", 5);
    P.rai = parse_rate(cfg.str("RATE_AI", "5Mb/s"));
    P.rhai = parse_rate(cfg.str("RATE_HAI", "50Mb/s"));
    P.min_rate = parse_rate(cfg.str("MIN_RATE", "100Mb/s"));
    P.dctcp_rai = parse_rate(cfg.str("DCTCP_RATE_AI", "1000Mb/s"));
    P.u_target = cfg.num("U_TARGET", 0.95);
    P.mi_thresh = cfg.u32("MI_THRESH", 5);
    P.fast_react = (int)cfg.u32("FAST_REACT", 1);
    P.sample_feedback = (int)cfg.u32("SAMPLE_FEEDBACK", 0);
    P.int_multi = cfg.u32("INT_MULTI", 1);
    P.buffer_size = cfg.u32("BUFFER_SIZE", 16) * 1024 * 1024;
    if (cfg.u32("GLOBAL_T", 1) != 1)
        return fail(TW_EINVAL, "GLOBAL_T 0 is not modelled for TwiNS topologies");
    auto ecn_lookup = [&](const char *key, uint64_t rate, double *v) {
        if (!cfg.has(key))
            return false;
        const vector<string> &x = config.at(key);
        int n = atoi(x[0].c_str());
        for (int i = 0; i < n && 2 + 2 * i < (int)x.size(); i++)
            if (strtoull(x[1 + 2 * i].c_str(), 0, 10) == rate)
            {
                *v = atof(x[2 + 2 * i].c_str());
                return true;
            }
        return false;
    };
    w.stop_ns = cfg.num("SIMULATOR_STOP_TIME", 10.0) * 1e9;

    // topology (src/rdma/rmain.cu:500-512): hosts, switches, links in file order
    int &H = w.H, &S = w.S;
    vector<std::pair<int, int>> links;
    {
        std::ifstream f(topo_path);
        if (!f)
            return fail(TW_EIO, "cannot open %s", topo_path);
        string x, a, b;
        while (f >> x)
        {
            if (x == "hosts")
                f >> H;
            else if (x == "switches")
                f >> S;
            else if (x == "link")
            {
                f >> a >> b;
                auto id = [&](const string &y) { return (y[0] == 'h' ? 0 : H) + atoi(y.c_str() + 1); };
                links.push_back({id(a), id(b)});
            }
            else if (isdigit((unsigned char)x[0]))
                return fail(TW_EINVAL, "%s: SimAI topologies are not modelled", topo_path);
        }
    }
    const int N = H + S, L = 2 * (int)links.size();
    if (H <= 0 || S < 0 || L == 0)
        return fail(TW_EINVAL, "%s: no hosts or links", topo_path);
    w.deg.assign(N, 0);
    for (auto &l : links)
    {
        if (l.first < 0 || l.first >= N || l.second < 0 || l.second >= N)
            return fail(TW_EINVAL, "%s: a link names node %d or %d of %d", topo_path, l.first, l.second, N);
        w.deg[l.first]++, w.deg[l.second]++;
    }
    int K = 1;
    for (int n = 0; n < N; n++)
        K = std::max(K, w.deg[n]);
    w.K = K;
    if (K > 255)
        return fail(TW_EINVAL, "a node has %d ports (at most 255)", K);
    for (int h = 0; h < H; h++)
        if (w.deg[h] != 1)
            return fail(TW_EINVAL, "host %d has %d links (one expected)", h, w.deg[h]);
    w.out_link.assign((size_t)N * K, -1);
    w.in_link.assign((size_t)N * K, -1);
    w.lbps.assign(L, bps);
    w.ldelay.assign(L, delay);
    vector<std::map<int, int>> nbr(N); // neighbour -> port (ns-3's nbr2if: by node id; the last link wins)
    {
        vector<int> d(N, 0);
        for (size_t i = 0; i < links.size(); i++)
        {
            int a = links[i].first, b = links[i].second, pa = d[a]++, pb = d[b]++;
            w.out_link[(size_t)a * K + pa] = 2 * (int)i;
            w.in_link[(size_t)a * K + pa] = 2 * (int)i + 1;
            w.out_link[(size_t)b * K + pb] = 2 * (int)i + 1;
            w.in_link[(size_t)b * K + pb] = 2 * (int)i;
            nbr[a][b] = pa;
            nbr[b][a] = pb;
        }
    }
    // routes: a BFS from every host over the switches, neighbours in id order; a switch's next
    // hops towards the host are the nodes one hop closer, in the order the BFS met them
    // (CalculateRoute, src/rdma/rmain.cu:572-660), with the path sums maxRtt and maxBdp need
    vector<vector<std::pair<int, int>>> adj(N);
    for (int n = 0; n < N; n++)
        adj[n].assign(nbr[n].begin(), nbr[n].end());
    vector<vector<int>> cur(S);
    vector<uint32_t> hoff((size_t)S * H);
    vector<uint8_t> hn((size_t)S * H), hports;
    vector<int32_t> hop_h((size_t)H * H, -1);
    vector<uint64_t> pdelay((size_t)H * H, 0), ptx((size_t)H * H, 0), pbw((size_t)H * H, 0);
    for (int h = 0; h < H; h++)
    {
        vector<int> q = {h}, dis(N, -1);
        vector<uint64_t> dl(N, 0), txd(N, 0), bw(N, 0);
        dis[h] = 0;
        bw[h] = ~0ull;
        for (size_t i = 0; i < q.size(); i++)
        {
            const int now = q[i], dd = dis[now];
            for (auto &e : adj[now])
            {
                const int next = e.first;
                if (dis[next] < 0)
                {
                    const int l = w.out_link[(size_t)now * K + e.second];
                    dis[next] = dd + 1;
                    dl[next] = dl[now] + (uint64_t)w.ldelay[l];
                    txd[next] = txd[now] + (uint64_t)P.mtu * 1000000000ull * 8 / w.lbps[l];
                    bw[next] = std::min(bw[now], w.lbps[l]);
                    if (next >= H)
                        q.push_back(next);
                }
                if (dd + 1 == dis[next] && next >= H)
                    cur[next - H].push_back(now);
            }
        }
        for (int i = 0; i < H; i++)
        {
            const size_t ih = (size_t)h * H + i;
            hop_h[ih] = dis[i];
            pdelay[ih] = dl[i], ptx[ih] = txd[i], pbw[ih] = bw[i];
        }
        for (int x = 0; x < S; x++)
        {
            hoff[(size_t)x * H + h] = (uint32_t)hports.size();
            hn[(size_t)x * H + h] = (uint8_t)cur[x].size();
            for (int y : cur[x])
                hports.push_back((uint8_t)nbr[H + x].at(y));
            cur[x].clear();
        }
    }
    w.rt.assign((size_t)S * H, rd::Route{});
    for (int x = 0; x < S; x++)
        for (int h = 0; h < H; h++)
        {
            rd::Route &r = w.rt[(size_t)x * H + h];
            r.off = (uint32_t)w.rtp.size();
            r.n = hn[(size_t)x * H + h];
            w.rtp.insert(w.rtp.end(), hports.begin() + hoff[(size_t)x * H + h], hports.begin() + hoff[(size_t)x * H + h] + r.n);
        }
    {
        int64_t far = 0;
        uint64_t bdp = 0;
        for (size_t i = 0; i < (size_t)H * H; i++)
            far = std::max<int64_t>(far, hop_h[i]);
        if (P.cc == 3 && far - 1 > rd::MAX_HOP)
            return fail(TW_EINVAL, "HPCC: paths of %lld switches (INT records at most %d)", (long long)(far - 1), rd::MAX_HOP);
        // maxRtt, maxBdp over host pairs (src/rdma/rmain.cu:661-690)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < H; i++)
            {
                const size_t ji = (size_t)j * H + i;
                const uint64_t rtt = pdelay[ji] * 2 + ptx[ji];
                w.max_rtt = std::max<uint64_t>(w.max_rtt, rtt);
                bdp = std::max<uint64_t>(bdp, rtt * pbw[ji] / 1000000000ull / 8);
            }
        w.max_bdp = (uint32_t)bdp;
    }
    const uint32_t win = cfg.u32("HAS_WIN", 1) ? w.max_bdp : 0;
    // switch ports: headroom, pfc_a_shift from the NIC rate (host 0's link), ECN thresholds by rate
    // (src/rdma/rmain.cu:692-721)
    {
        const uint64_t nic_rate = w.lbps[w.out_link[0]];
        for (int x = 0; x < S; x++)
            for (int p = 0; p < w.deg[H + x]; p++)
            {
                const int l = w.out_link[(size_t)(H + x) * K + p];
                rd::PortCfg pc = {};
                uint64_t r = w.lbps[l];
                pc.bps = r;
                pc.hdrm_cap = (uint32_t)(r * (uint64_t)w.ldelay[l] / 8 / 1000000000ull * 3);
                pc.shift = 3;
                while (r > nic_rate && pc.shift > 0)
                {
                    pc.shift--;
                    r /= 2;
                }
                double kmin = 0, kmax = 0, pmax = 0;
                if (!ecn_lookup("KMIN_MAP", pc.bps, &kmin) || !ecn_lookup("KMAX_MAP", pc.bps, &kmax) || !ecn_lookup("PMAX_MAP", pc.bps, &pmax))
                    return fail(TW_EINVAL, "KMIN/KMAX/PMAX_MAP: no entry for %llu bps", (unsigned long long)pc.bps);
                pc.kmin = (uint32_t)kmin * 1000;
                pc.kmax = (uint32_t)kmax * 1000;
                pc.pmax = pmax;
                pc.link_in = w.in_link[(size_t)(H + x) * K + p];
                pc.link_out = l;
                w.pcfg.push_back(pc);
            }
    }
    // flows: rows in file order, then the start order (stable by start), source ports per host pair
    // from 10000 and the StartApplication chains (src/rdma/rmain.cu:844-911)
    {
        std::ifstream f(flow_path);
        if (!f)
            return fail(TW_EIO, "cannot open %s", flow_path);
        string line;
        std::getline(f, line);
        while (std::getline(f, line))
        {
            unsigned a, b;
            unsigned long long sz, st;
            if (sscanf(line.c_str(), "%u,%u,%llu,%llu", &a, &b, &sz, &st) != 4)
                continue;
            if ((int)a >= H || (int)b >= H)
                return fail(TW_EINVAL, "a flow between %u and %u: not hosts", a, b);
            rd::FlowInfo x = {};
            x.src = (int)a, x.dst = (int)b, x.size = sz, x.start = (int64_t)st;
            x.win = win;
            x.base_rtt = w.max_rtt;
            w.fi.push_back(x);
        }
    }
    const int F = (int)w.fi.size();
    vector<int> order(F);
    for (int i = 0; i < F; i++)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w.fi[a].start < w.fi[b].start; });
    {
        std::map<std::pair<int, int>, uint32_t> port;
        int64_t prev_t = 0;
        int prev_n = 0, k = 0;
        for (int i = 0; i < F;)
        {
            int j = i;
            while (j < F && w.fi[order[j]].start == w.fi[order[i]].start)
                j++;
            const int64_t t = w.fi[order[i]].start;
            for (int m = i; m < j; m++)
            {
                rd::FlowInfo &x = w.fi[order[m]];
                auto key = std::make_pair(x.src, x.dst);
                if (!port.count(key))
                    port[key] = 10000;
                x.sport = port[key]++;
                rd::Chain c = {};
                c.root = (uint32_t)m + 1;
                host_push(&c, t - prev_t, k == 0 ? 0 : prev_n); // ScheduleFlowInputs
                host_push(&c, 0, m - i);                        // Application::Start
                host_push(&c, 0, 0);                            // StartApplication
                x.start_chain = c;                              // (hash and ancestors 0: they cover what follows)
            }
            prev_t = t;
            prev_n = j - i;
            k++;
            i = j;
        }
    }
    w.start_off.assign(H + 1, 0);
    for (int i = 0; i < F; i++)
        w.start_off[w.fi[order[i]].src + 1]++;
    for (int h = 0; h < H; h++)
        w.start_off[h + 1] += w.start_off[h];
    w.start_list.assign(F, 0);
    {
        vector<int> pos(w.start_off.begin(), w.start_off.end() - 1);
        for (int i = 0; i < F; i++)
            w.start_list[pos[w.fi[order[i]].src]++] = order[i];
    }
    *out = wp.release();
    return TW_OK;
}

extern "C" int trd_world_info(const trd_world *w, int32_t *hosts, int32_t *switches, int32_t *flows, uint64_t *max_rtt_ns, uint32_t *max_bdp)
{
    if (!w)
        return fail(TW_EINVAL, "trd_world_info: null world");
    if (hosts)
        *hosts = w->H;
    if (switches)
        *switches = w->S;
    if (flows)
        *flows = (int32_t)w->fi.size();
    if (max_rtt_ns)
        *max_rtt_ns = w->max_rtt;
    if (max_bdp)
        *max_bdp = w->max_bdp;
    return TW_OK;
}
extern "C" void trd_world_destroy(trd_world *w) { delete w; }

// ------------------------------------------------------------------ the batch on the device
struct trd_batch
{
    rd::Sim s = {};     // device pointers
    rd::Sim *sim_d = nullptr;
    const rd::Sim **sims_d = nullptr; // [1] sim_d, for the launch configuration
    vector<void *> allocs;
    Arena *arena = nullptr; // a server's batch: its memory comes from the server's arena
    bool own_stream = true;
    int W = 0, H = 0, S = 0, F = 0;
    vector<int> fbase, nf;
    vector<int64_t> start; // [F] flows' starts, by global id
    int threads = 256;
    int team_blocks = 0; // 0: from the worlds' sizes
    int hpw = 0;         // hosts per host item (0: from the worlds' sizes)
    vector<int> hs, ss;  // per world: hosts, switches
    int *queue_d = nullptr;
    unsigned *qnext_d = nullptr;
    rd::TeamState *teams_d = nullptr;
    rd::LaunchCfg *cfg_d = nullptr;
    rd::SwWorkWide *wide_d = nullptr; // a switch workspace per wave of a launch (worlds with switches of KS + 1 to KW ports)
    int max_teams = 0;
    int kmax = 0; // the most ports of a switch
    int tail_lg = 0; // the largest world's tail ring: 2^tail_lg nodes
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ran = false;
    vector<rd::WorldRun> runs;
    ~trd_batch()
    {
        for (void *p : allocs)
            if (arena)
                arena->release(p);
            else
                (void)hipFree(p);
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
        if (st && own_stream)
            (void)hipStreamDestroy(st);
    }
    template <typename T>
    int alloc(T **p, size_t n)
    {
        void *d = nullptr;
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (arena)
        {
            if (!(d = arena->alloc(bytes)))
                return fail(TW_ENOMEM, "the device arena has no %zu free bytes", bytes);
        }
        else if (hipMalloc(&d, bytes) != hipSuccess)
            return fail(TW_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
        allocs.push_back(d);
        if (hipMemsetAsync(d, 0, bytes, st) != hipSuccess)
            return fail(TW_EDEVICE, "hipMemset failed");
        *p = (T *)d;
        return TW_OK;
    }
    // host copies of the uploads, kept until the stream has taken them (one synchronization at the
    // end of the packing, not one per array: a serving loop's packing waits less for the device)
    vector<vector<char>> staged;
    template <typename T>
    int upload(T **p, const vector<T> &v)
    {
        int e = alloc(p, v.size());
        if (e)
            return e;
        if (v.empty())
            return TW_OK;
        staged.emplace_back((const char *)v.data(), (const char *)v.data() + v.size() * sizeof(T));
        if (hipMemcpyAsync(*p, staged.back().data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(TW_EDEVICE, "hipMemcpy failed");
        return TW_OK;
    }
    int sync_uploads()
    {
        if (hipStreamSynchronize(st) != hipSuccess)
            return fail(TW_EDEVICE, "hipStreamSynchronize failed");
        staged.clear();
        return TW_OK;
    }
    template <typename T>
    int upload_c(const T **p, const vector<T> &v)
    {
        T *q = nullptr;
        int e = upload(&q, v);
        *p = q;
        return e;
    }
};

#define TRY(x)                                                                                                                   \
    do                                                                                                                            \
    {                                                                                                                             \
        int e_ = (x);                                                                                                             \
        if (e_)                                                                                                                   \
            return e_;                                                                                                            \
    } while (0)

static int pack(trd_batch &b, trd_world *const *worlds, int W, const trd_pack_cfg &cfg)
{
    rd::Sim &s = b.s;
    // the slot: the shortest link delay of all worlds (src/rdma/rmain.cu:1253-1266)
    int64_t D = INT64_MAX;
    for (int i = 0; i < W; i++)
        for (int64_t d : worlds[i]->ldelay)
            D = std::min(D, d);
    for (int i = 0; i < W; i++)
    {
        const rd::Params &P = worlds[i]->P;
        if (P.cc == 1 && (P.alpha_int < D || P.dec_int < D || P.rp_int < D))
            return fail(TW_EINVAL, "world %d: DCQCN timer intervals must be at least the slot (%lld ns)", i, (long long)D);
    }
    int H = 0, S = 0, L = 0, F = 0, NP = 0;
    vector<int> hbase(W), sbase(W), lbase(W), fbase(W), pbase0(W);
    for (int i = 0; i < W; i++)
    {
        const trd_world &w = *worlds[i];
        hbase[i] = H, sbase[i] = S, lbase[i] = L, fbase[i] = F, pbase0[i] = NP;
        H += w.H, S += w.S, L += w.L(), F += (int)w.fi.size(), NP += (int)w.pcfg.size();
    }
    if (H > 65535)
        return fail(TW_EINVAL, "too many hosts (%d; frames carry 16-bit host ids)", H);
    b.W = W, b.H = H, b.S = S, b.F = F;
    b.fbase = fbase;
    s.H = H, s.S = S, s.F = F, s.W = W, s.L = L, s.D = D;
    s.serial = cfg.serial ? 1 : 0;
    // the batch's stop: the latest world's (src/rdma/rmain.cu:1290)
    int64_t max_slots = 0;
    for (int i = 0; i < W; i++)
        max_slots = std::max(max_slots, (int64_t)std::min(worlds[i]->stop_ns / (double)D, 1e15));
    uint32_t mtu = 0;
    for (int i = 0; i < W; i++)
        mtu = std::max(mtu, worlds[i]->P.mtu);

    vector<rd::Params> Pw(W);
    vector<rd::World> wd(W);
    vector<int> h_world(H), h_lin(H), h_lout(H), s_world(S), start_list(F), start_off(H + 1, 0);
    vector<rd::SwInfo> si(S);
    vector<rd::PortCfg> pc(NP);
    vector<rd::Route> rt;
    vector<uint8_t> rtp;
    vector<rd::FlowInfo> fi(F);
    vector<rd::Link> li(L);
    vector<uint64_t> rates;
    size_t nframes = 0, ncnt = 0;
    int pend_total = 0;
    vector<rd::Chain> root_chain(F);
    for (int wi = 0; wi < W; wi++)
    {
        const trd_world &w = *worlds[wi];
        Pw[wi] = w.P;
        rd::World &z = wd[wi];
        z.h0 = hbase[wi], z.nh = w.H, z.s0 = sbase[wi], z.ns = w.S, z.f0 = fbase[wi], z.nf = (int)w.fi.size();
        z.max_slots = max_slots;
        for (int h = 0; h < w.H; h++)
        {
            h_world[hbase[wi] + h] = wi;
            h_lin[hbase[wi] + h] = lbase[wi] + w.in_link[(size_t)h * w.K];
            h_lout[hbase[wi] + h] = lbase[wi] + w.out_link[(size_t)h * w.K];
            start_off[hbase[wi] + h + 1] = w.start_off[h + 1] - w.start_off[h];
        }
        const uint32_t rb = (uint32_t)rtp.size();
        rtp.insert(rtp.end(), w.rtp.begin(), w.rtp.end());
        int pb = pbase0[wi];
        for (int x = 0; x < w.S; x++)
        {
            rd::SwInfo &q = si[sbase[wi] + x];
            q.seed = (uint32_t)(w.H + x);
            q.hbase = hbase[wi];
            q.rt_off = (uint32_t)rt.size();
            q.pbase = pb;
            q.deg = w.deg[w.H + x];
            b.kmax = std::max(b.kmax, (int)q.deg);
            for (int h = 0; h < w.H; h++)
            {
                rd::Route r = w.rt[(size_t)x * w.H + h];
                r.off += rb;
                rt.push_back(r);
            }
            s_world[sbase[wi] + x] = wi;
            pb += q.deg;
        }
        for (size_t p = 0; p < w.pcfg.size(); p++)
        {
            rd::PortCfg c = w.pcfg[p];
            c.link_in += lbase[wi];
            c.link_out += lbase[wi];
            pc[pbase0[wi] + p] = c;
        }
        for (size_t f = 0; f < w.fi.size(); f++)
        {
            rd::FlowInfo x = w.fi[f];
            x.start_chain.root += (uint32_t)fbase[wi]; // roots over all worlds (the order within one kept)
            root_chain[x.start_chain.root - 1] = x.start_chain;
            x.src += hbase[wi];
            x.dst += hbase[wi];
            fi[fbase[wi] + f] = x;
        }
        for (size_t f = 0; f < w.start_list.size(); f++)
            start_list[fbase[wi] + f] = fbase[wi] + w.start_list[f];
        // links: rate classes and the rings of arrival slots; a frame started in slot k arrives
        // by slot k + 1 + (txmax + delay - 1) / D; a bucket holds the frames of one slot at one
        // per shortest frame (src/rdma/rmain.cu:1800-1855)
        int pend_n = 2;
        for (int l = 0; l < w.L(); l++)
        {
            rd::Link &x = li[lbase[wi] + l];
            x.delay = w.ldelay[l];
            x.bps = w.lbps[l];
            size_t r = std::find(rates.begin(), rates.end(), x.bps) - rates.begin();
            if (r == rates.size())
                rates.push_back(x.bps);
            x.rc = (uint32_t)r;
            const int64_t txmax = rd::tx_ns(x.bps, mtu + 128), txmin = rd::tx_ns(x.bps, 60);
            const int64_t nb = 2 + (txmax + x.delay - 1) / D;
            const int64_t cap = txmin == 0 ? rd::MAXF : std::min<int64_t>(rd::MAXF, D / txmin + 2);
            if (nb > 4096)
                return fail(TW_EINVAL, "a link's delay (%lld ns) is too long for the slot (%lld ns)", (long long)x.delay, (long long)D);
            x.nb = (uint32_t)nb;
            x.cap = (uint32_t)cap;
            x.moff = nframes;
            x.coff = (uint32_t)ncnt;
            nframes += (size_t)nb * cap;
            ncnt += nb;
            pend_n = std::max(pend_n, (int)nb);
        }
        z.pend_off = pend_total;
        z.pend_n = pend_n;
        pend_total += pend_n;
    }
    for (int h = 0; h < H; h++)
        start_off[h + 1] += start_off[h];
    // the roots in the walk's order of their start chains: first differing lead (larger first), the
    // shorter chain, then the root (src/rdma/rmain.cu:1541-1569)
    vector<uint32_t> rank(F, 0);
    {
        auto leads = [](const rd::Chain &c) {
            vector<uint32_t> v;
            for (int i = 0; i < rd::RUNS && c.r[i].code; i++)
                for (uint32_t k = 0; k < c.r[i].cnt; k++)
                    v.push_back(c.r[i].code >> rd::ORD_BITS);
            return v;
        };
        vector<vector<uint32_t>> Ls(F);
        vector<int> idx(F);
        for (int i = 0; i < F; i++)
        {
            Ls[i] = leads(root_chain[i]);
            idx[i] = i;
        }
        std::sort(idx.begin(), idx.end(), [&](int a, int c) {
            const vector<uint32_t> &x = Ls[a], &y = Ls[c];
            for (size_t i = 0; i < x.size() && i < y.size(); i++)
                if (x[i] != y[i])
                    return x[i] > y[i];
            if (x.size() != y.size())
                return x.size() < y.size();
            return root_chain[a].root < root_chain[c].root;
        });
        for (int i = 0; i < F; i++)
            rank[idx[i]] = (uint32_t)i;
    }
    TRY(b.upload_c(&s.Pw, Pw));
    TRY(b.upload_c(&s.wd, wd));
    {
        vector<rd::WorldRun> runs(W);
        for (auto &r : runs)
            r.next_wake = ~0ull;
        TRY(b.upload(&s.run, runs));
    }
    TRY(b.upload_c(&s.h_world, h_world));
    TRY(b.upload_c(&s.h_link_in, h_lin));
    TRY(b.upload_c(&s.h_link_out, h_lout));
    TRY(b.upload_c(&s.si, si));
    TRY(b.upload_c(&s.s_world, s_world));
    TRY(b.upload_c(&s.pc, pc));
    TRY(b.upload_c(&s.rt, rt));
    TRY(b.upload_c(&s.rtp, rtp));
    TRY(b.upload_c(&s.fi, fi));
    TRY(b.upload_c(&s.root_rank, rank));
    TRY(b.upload_c(&s.start_list, start_list));
    TRY(b.alloc(&s.qp, F));
    TRY(b.alloc(&s.qc, F));
    TRY(b.alloc(&s.rx, F));
    {
        vector<rd::Host> hs(H);
        for (int h = 0; h < H; h++)
        {
            hs[h] = rd::Host{};
            hs[h].tc_at = hs[h].ns_at = -1;
            hs[h].ob_last = -(1ll << 16);
            hs[h].due_min = INT64_MAX;
            hs[h].next_start = start_off[h];
            hs[h].start_end = start_off[h + 1];
            hs[h].ta.wbits = (uint64_t)h_world[h] << rd::TAIL_WSHIFT;
        }
        TRY(b.upload(&s.hs, hs));
    }
    TRY(b.alloc(&s.h_tc, H));
    TRY(b.alloc(&s.h_ns, H));
    TRY(b.alloc(&s.ackq, (size_t)H * rd::ACKQ_CAP));
    {
        vector<int> per(H, 0);
        for (auto &x : fi)
            per[x.src]++;
        s.qps_cap = std::max(1, *std::max_element(per.begin(), per.end()));
    }
    TRY(b.alloc(&s.qps, (size_t)H * s.qps_cap));
    TRY(b.alloc(&s.due, (size_t)H * rd::MAX_DUE));
    {
        vector<rd::Switch> sw(S);
        vector<rd::Port> ps(NP);
        for (int x = 0; x < S; x++)
        {
            const rd::SwInfo &q = si[x];
            sw[x] = rd::Switch{};
            for (int p = 0; p < q.deg; p++)
            {
                sw[x].total_hdrm += pc[q.pbase + p].hdrm_cap;
                sw[x].max_shift = std::max(sw[x].max_shift, pc[q.pbase + p].shift);
            }
            sw[x].total_rsrv = rd::RESERVE * (uint32_t)q.deg;
            for (int p = 0; p < q.deg; p++)
            {
                rd::Port &pt = ps[q.pbase + p];
                pt = rd::Port{};
                pt.ob_last = -(1ll << 16);
                pt.tc_at = -1;
                pt.ta.wbits = (uint64_t)s_world[x] << rd::TAIL_WSHIFT;
                // ConfigEcnStream(port, node * SwitchMmu::pCnt + device)
                mrg_stream6(1, (uint64_t)q.seed * 257 + (uint64_t)(p + 1), 1, pt.rng);
            }
        }
        TRY(b.upload(&s.sw, sw));
        TRY(b.upload(&s.ps, ps));
        vector<int> busy(S, 1);
        TRY(b.upload(&s.sw_busy, busy));
    }
    TRY(b.alloc(&s.p_tc, NP));
    TRY(b.alloc(&s.pr, NP));
    s.q0_cap = 64;
    s.q3_cap = cfg.q3cap > 0 ? cfg.q3cap : 4096;
    TRY(b.alloc(&s.q0, (size_t)NP * s.q0_cap));
    TRY(b.alloc(&s.q3, (size_t)NP * s.q3_cap));
    TRY(b.upload_c(&s.li, li));
    TRY(b.alloc(&s.msg, nframes));
    // (16-byte coherent frame stores address the rings with 32-bit offsets)
    s.msg_bytes = (uint64_t)nframes * sizeof(rd::LFrame) <= 0x7FFFFFF0ull ? (uint32_t)((uint64_t)nframes * sizeof(rd::LFrame)) : 0;
    TRY(b.upload(&s.msg_cnt, vector<int64_t>(ncnt, -(1ll << 16))));
    {
        vector<int32_t> txt(rates.size() * rd::TXT_N);
        for (size_t r = 0; r < rates.size(); r++)
            for (int n = 0; n < rd::TXT_N; n++)
                txt[r * rd::TXT_N + n] = (int32_t)rd::tx_ns(rates[r], (uint32_t)n);
        TRY(b.upload_c(&s.txt, txt));
        s.R = (int)rates.size();
    }
    TRY(b.alloc(&s.pend, (size_t)pend_total));
    {
        // a ring of tail nodes per world: 2^(log2(entities) + 14) nodes, from 2^16 to 2^tail_log2;
        // with tail_exact, 2^tail_log2 nodes each
        const int lgmax = cfg.tail_log2 > 0 ? std::min(cfg.tail_log2, 30) : 24;
        if (cfg.tail_exact && (cfg.tail_log2 < 4 || cfg.tail_log2 > 30))
            return fail(TW_EINVAL, "tail_log2 %d: from 4 to 30", cfg.tail_log2);
        vector<rd::TailRing> rings(W);
        uint64_t total = 0;
        b.tail_lg = 0;
        for (int wi = 0; wi < W; wi++)
        {
            int lg = 0;
            while ((1 << lg) < wd[wi].nh + wd[wi].ns)
                lg++;
            lg = cfg.tail_exact ? cfg.tail_log2 : std::max(16, std::min(lgmax, lg + 14));
            b.tail_lg = std::max(b.tail_lg, lg);
            rings[wi].off = total;
            rings[wi].mask = ((uint64_t)1 << lg) - 1;
            total += (uint64_t)1 << lg;
        }
        TRY(b.alloc(&s.tail, (size_t)total));
        TRY(b.upload_c(&s.rings, rings));
        TRY(b.alloc(&s.tail_cnt, W));
        const uint64_t mcap = 1ull << 22;
        TRY(b.alloc(&s.memo, mcap));
        s.memo_mask = mcap - 1;
    }
    TRY(b.upload(&s.fct, vector<int64_t>(F, -1)));
    TRY(b.alloc(&s.ev, rd::NEV));
    TRY(b.upload(&b.sim_d, vector<rd::Sim>{s}));
    b.start.resize(F);
    for (int f = 0; f < F; f++)
        b.start[f] = fi[f].start;
    b.nf.resize(W);
    b.hs.resize(W);
    b.ss.resize(W);
    for (int wi = 0; wi < W; wi++)
    {
        b.nf[wi] = (int)worlds[wi]->fi.size();
        b.hs[wi] = worlds[wi]->H;
        b.ss[wi] = worlds[wi]->S;
    }
    TRY(b.upload(&b.sims_d, vector<const rd::Sim *>{b.sim_d}));
    TRY(b.sync_uploads());
    if (b.arena)
        return TW_OK; // (a server's batch: the server launches it)
    // launches: the queue of worlds, the teams' barrier state
    const int cap = rd::launch_capacity(b.threads);
    if (cap <= 0)
        return fail(TW_EDEVICE, "the rdma kernel fits no workgroup on the GPU");
    b.max_teams = cap;
    TRY(b.alloc(&b.queue_d, W));
    TRY(b.alloc(&b.qnext_d, 1));
    TRY(b.alloc(&b.teams_d, cap));
    TRY(b.alloc(&b.cfg_d, 1));
    if (b.kmax > rd::KS && b.kmax <= rd::KW)
        TRY(b.alloc(&b.wide_d, (size_t)cap * rd::LAUNCH_WAVES));
    return TW_OK;
}

extern "C" int trd_batch_create(trd_world *const *worlds, int32_t n, const trd_pack_cfg *cfg, trd_batch **out)
{
    if (!worlds || n <= 0 || !out)
        return fail(TW_EINVAL, "trd_batch_create: no worlds");
    for (int i = 0; i < n; i++)
        if (!worlds[i])
            return fail(TW_EINVAL, "trd_batch_create: world %d is null", i);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(TW_EDEVICE, "no GPU: the RDMA model runs on the GPU only");
    trd_pack_cfg c = cfg ? *cfg : trd_pack_cfg{};
    std::unique_ptr<trd_batch> b(new trd_batch);
    b->threads = 256;
    b->team_blocks = c.team_blocks;
    b->hpw = c.hosts_per_wave > 0 ? std::min(c.hosts_per_wave, 64) : 0;
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->e0) != hipSuccess ||
        hipEventCreate(&b->e1) != hipSuccess)
        return fail(TW_EDEVICE, "hipStreamCreate failed");
    TRY(pack(*b, worlds, n, c));
    *out = b.release();
    return TW_OK;
}

extern "C" int trd_batch_run(trd_batch *b, trd_run_stats *st)
{
    if (!b)
        return fail(TW_EINVAL, "trd_batch_run: null batch");
    if (b->ran)
        return fail(TW_ESTATE, "trd_batch_run: the batch has run");
    b->ran = true;
    const auto t0 = std::chrono::steady_clock::now();
    b->runs.assign(b->W, rd::WorldRun{});
    int launches = 0;
    double kernel_s = 0;
    int64_t prev = -1;
    // the launch shape: hosts per host item (fewer: less divergence among a wave's hosts, more
    // waves), and one team size. Batches of up to 8 worlds take the fewest hosts per item (2, 4, 8
    // or 16) whose items still fit a wave each of the largest team - 128 workgroups for a world
    // alone, the GPU's share of each of up to 8 worlds (at least 32) - and a workgroup per item of
    // the largest world up to that team. Measured with the slot barrier without cache fences:
    // FatTree k=8 a2a50s (HPCC) alone 1.14, 0.80 and 0.70 s on 24 workgroups at 8 hosts per item, 64
    // at 4 and 128 at 2; k8_perm 0.196 and 0.132 s on 24 at 8 and 96 at 2; Jellyfish a2a30 1.44 and
    // 1.20 s on 12 at 8 and 96 at 2; FatTree k=16 a2a30 (DCQCN) 1.51 s on 112 at 8 and on 192 or 256
    // alike; the 512-host leaf-spine (HPCC + DCQCN side by side) 1.27 s of kernel on 22 at 8 (a wave
    // per item), 1.15 s on 48 at 4 and on 128 at 2. Larger batches: 64 hosts per item, teams of 16
    // workgroups with a wave per item. As many teams as the GPU holds at once.
    const int cap = b->max_teams;
    const bool small = b->W <= 8;
    const int most = b->W == 1 ? std::min(128, cap) : small ? std::max(32, cap / b->W) : 16;
    int hpw = b->hpw;
    if (hpw <= 0)
    {
        hpw = small ? 2 : 64;
        for (bool fits = false; small && !fits && hpw < 16;)
        {
            fits = true;
            for (int i = 0; i < b->W; i++)
                fits = fits && (b->hs[i] + hpw - 1) / hpw + b->ss[i] <= most * 4;
            if (!fits)
                hpw *= 2;
        }
    }
    vector<int> items(b->W);
    for (int i = 0; i < b->W; i++)
        items[i] = (b->hs[i] + hpw - 1) / hpw + b->ss[i];
    int T = b->team_blocks;
    if (T <= 0)
    {
        int mx = 1;
        for (int v : items)
            mx = std::max(mx, v);
        T = std::min(most, small ? mx : (mx + 3) / 4);
    }
    T = std::max(1, std::min(T, cap));
    // worlds by work items, largest first
    vector<int> order(b->W);
    for (int i = 0; i < b->W; i++)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return items[a] > items[c]; });
    for (;;)
    {
        vector<int> queue;
        for (int w : order)
            if (!b->runs[w].finished)
                queue.push_back(w);
        const int nteams = std::min<int>((int)queue.size(), cap / T);
        rd::LaunchCfg L = {};
        L.T = T;
        L.hpw = hpw;
        L.nq = (int)queue.size();
        L.queue = b->queue_d;
        L.qnext = b->qnext_d;
        L.teams = b->teams_d;
        L.sims = b->sims_d; // (queue entries: batch 0's worlds)
        L.wide = b->wide_d;
        if (hipMemcpyAsync(b->queue_d, queue.data(), sizeof(int) * queue.size(), hipMemcpyHostToDevice, b->st) != hipSuccess ||
            hipMemcpyAsync(b->cfg_d, &L, sizeof L, hipMemcpyHostToDevice, b->st) != hipSuccess ||
            hipMemsetAsync(b->qnext_d, 0, sizeof(unsigned), b->st) != hipSuccess ||
            hipMemsetAsync(b->teams_d, 0, sizeof(rd::TeamState) * nteams, b->st) != hipSuccess)
            return fail(TW_EDEVICE, "launch set-up failed");
        (void)hipEventRecord(b->e0, b->st);
        int e = rd::launch_worlds(b->cfg_d, nteams * T, b->threads, b->st);
        if (e)
            return fail(TW_EDEVICE, "rdma kernel launch failed (%d)", e);
        (void)hipEventRecord(b->e1, b->st);
        launches++;
        if (hipMemcpyAsync(b->runs.data(), b->s.run, sizeof(rd::WorldRun) * b->W, hipMemcpyDeviceToHost, b->st) != hipSuccess ||
            hipStreamSynchronize(b->st) != hipSuccess)
            return fail(TW_EDEVICE, "the rdma kernel failed: %s", hipGetErrorString(hipGetLastError()));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, b->e0, b->e1);
        kernel_s += ms * 1e-3;
        bool all = true;
        int64_t total = 0;
        for (auto &r : b->runs)
        {
            all &= r.finished != 0;
            total += r.slots_run;
        }
        {
            vector<rd::TeamState> ts(nteams);
            if (nteams && hipMemcpy(ts.data(), b->teams_d, sizeof(rd::TeamState) * nteams, hipMemcpyDeviceToHost) != hipSuccess)
                return fail(TW_EDEVICE, "hipMemcpy failed");
            for (auto &t : ts)
                if (t.err)
                    return fail(TW_EDEVICE, "a team barrier timed out");
        }
        if (all)
            break;
        if (total == prev)
            return fail(TW_EDEVICE, "a launch ran no slot");
        prev = total;
    }
    if (st)
    {
        *st = trd_run_stats{};
        for (auto &r : b->runs)
        {
            st->slots = std::max<int64_t>(st->slots, r.slots_run);
            st->sim_ns = std::max<int64_t>(st->sim_ns, r.slot * b->s.D);
            st->flows_done += r.flows_done;
        }
        st->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st->kernel_s = kernel_s;
        st->launches = launches;
    }
    return TW_OK;
}

extern "C" int trd_fct_get(const trd_batch *b, int32_t world, int64_t *start_ns, int64_t *fct_ns)
{
    if (!b || world < 0 || world >= b->W)
        return fail(TW_EINVAL, "trd_fct_get: bad world %d", world);
    const int f0 = b->fbase[world], n = b->nf[world];
    if (fct_ns && n > 0 && hipMemcpy(fct_ns, b->s.fct + f0, sizeof(int64_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(TW_EDEVICE, "hipMemcpy failed");
    if (start_ns)
        for (int i = 0; i < n; i++)
            start_ns[i] = b->start[f0 + i];
    return TW_OK;
}
extern "C" int trd_counters_get(const trd_batch *b, uint64_t *ev40)
{
    if (!b || !ev40)
        return fail(TW_EINVAL, "trd_counters_get: null argument");
    if (hipMemcpy(ev40, b->s.ev, sizeof(uint64_t) * rd::NEV, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(TW_EDEVICE, "hipMemcpy failed");
    return TW_OK;
}
// ------------------------------------------------------------------ the transmission trace
extern "C" int trd_txtrace_enable(trd_batch *b, int64_t from_ns, int64_t until_ns, uint64_t max_records)
{
    if (!b || b->arena)
        return fail(TW_EINVAL, "txtrace: a batch of its own");
    if (b->s.txt_rec.rec)
        return fail(TW_EINVAL, "txtrace: already enabled");
    if (b->ran)
        return fail(TW_EINVAL, "txtrace: enable it before trd_batch_run");
    rd::TxTrace t{};
    t.cap = max_records ? max_records : 64ull << 20;
    t.from = from_ns;
    t.until = until_ns;
    TRY(b->alloc(&t.rec, t.cap));
    TRY(b->alloc(&t.cnt, 1)); // (zeroed on the batch's stream, ahead of its launches)
    if (hipMemcpy((char *)b->sim_d + offsetof(rd::Sim, txt_rec), &t, sizeof t, hipMemcpyHostToDevice) != hipSuccess)
        return fail(TW_EDEVICE, "txtrace: device set-up failed");
    b->s.txt_rec = t;
    return TW_OK;
}

extern "C" int trd_txtrace_write(const trd_batch *b, int32_t world, const char *path, uint64_t *records, uint64_t *seen)
{
    if (!b || !path || world < 0 || world >= b->W)
        return fail(TW_EINVAL, "txtrace: bad batch, world or path");
    const rd::TxTrace &t = b->s.txt_rec;
    if (!t.rec)
        return fail(TW_EINVAL, "txtrace: not enabled (trd_txtrace_enable)");
    unsigned long long n = 0;
    if (hipMemcpy(&n, t.cnt, sizeof n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(TW_EDEVICE, "hipMemcpy failed");
    if (seen)
        *seen = n;
    n = std::min<unsigned long long>(n, t.cap);
    vector<rd::TxRec> r(n);
    if (n && hipMemcpy(r.data(), t.rec, n * sizeof(rd::TxRec), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(TW_EDEVICE, "hipMemcpy failed");
    // this world's records in its own numbering: hosts 0..H-1, switches H.. (src/rdma/rmain.cu:2262-2282)
    int h0 = 0, s0 = 0;
    for (int i = 0; i < world; i++)
        h0 += b->hs[i], s0 += b->ss[i];
    const int nh = b->hs[world], ns = b->ss[world];
    vector<rd::TxRec> mine;
    for (rd::TxRec x : r)
    {
        int node;
        if (x.node < b->H)
        {
            if (x.node < h0 || x.node >= h0 + nh)
                continue;
            node = x.node - h0;
        }
        else
        {
            const int sw = x.node - b->H;
            if (sw < s0 || sw >= s0 + ns)
                continue;
            node = nh + sw - s0;
        }
        x.node = node;
        if (x.kind < rd::K_PAUSE)
            x.a -= (uint32_t)h0, x.b -= (uint32_t)h0;
        mine.push_back(x);
    }
    // a link starts one transmission at a time: (time, node, device) orders them totally
    std::sort(mine.begin(), mine.end(), [](const rd::TxRec &a, const rd::TxRec &b) {
        return a.ts != b.ts ? a.ts < b.ts : a.node != b.node ? a.node < b.node : a.dev < b.dev;
    });
    FILE *f = fopen(path, "w");
    if (!f)
        return fail(TW_EIO, "cannot write %s", path);
    const char kinds[] = "DANPR";
    for (const rd::TxRec &x : mine)
        fprintf(f, "%lld %d %d %c %u %u %u %u %u %u\n", (long long)x.ts, x.node, x.dev, x.kind < 5 ? kinds[x.kind] : '?', x.a, x.b, x.sport,
                x.seq, (unsigned)x.size, (unsigned)x.ecn);
    if (fclose(f))
        return fail(TW_EIO, "cannot write %s", path);
    if (records)
        *records = mine.size();
    return TW_OK;
}

extern "C" int trd_batch_tail_log2(const trd_batch *b, int32_t *log2_nodes)
{
    if (!b || !log2_nodes)
        return fail(TW_EINVAL, "trd_batch_tail_log2: null argument");
    *log2_nodes = b->tail_lg;
    return TW_OK;
}

extern "C" int trd_world_stats_get(const trd_batch *b, int32_t world, int64_t *slots, double *device_s)
{
    if (!b || world < 0 || world >= b->W || b->runs.empty())
        return fail(TW_EINVAL, "trd_world_stats_get: bad world %d or no run", world);
    if (slots)
        *slots = b->runs[world].slots_run;
    if (device_s)
        *device_s = b->runs[world].ticks * 1e-8; // the 100 MHz constant clock
    return TW_OK;
}
extern "C" void trd_batch_destroy(trd_batch *b) { delete b; }

// ------------------------------------------------------------------ the serving loop (RDMA requests)
// Requests are worlds; each is packed into the server's device arena as a batch of its own (no
// hipMalloc or hipFree while kernels run, so a TCP serving loop's persistent kernel is never
// waited for), and a launch of the team kernel runs resident, unfinished requests' worlds
// (LaunchCfg::sims: one Sim per request). Worlds of two sizes launch on two streams at once: small
// worlds on small teams, large ones on large teams, each class within its share of max_blocks
// workgroups (by default a CU in eight, which a TCP serving loop can leave free: then both loops'
// teams always fit on the GPU together). A launch ends after 1 s of device time; requests that
// arrived meanwhile join its class's next launch.
constexpr int RSRV_TAIL_MAX = 26; // the largest tail ring a rerun of a request gets (2^26 nodes)
struct RSrvReq
{
    std::unique_ptr<trd_world> w;
    trd_batch *b = nullptr;
    int state = TW_REQ_WAITING;
    int items = 0;   // work items per slot
    int tail_lg = 0; // 0: the server's automatic tail rings; else 2^tail_lg nodes (a rerun after lost tails)
};
// one class's launches: its stream, its configuration in device memory and host copies (kept until
// the launch ends)
struct RSrvLane
{
    hipStream_t ks = nullptr;
    hipEvent_t done_ev = nullptr;
    bool in_flight = false;
    int blocks = 0;
    int *queue_d = nullptr;
    unsigned *qnext_d = nullptr;
    rd::TeamState *teams_d = nullptr;
    rd::LaunchCfg *cfg_d = nullptr;
    const rd::Sim **sims_d = nullptr;
    rd::SwWorkWide *wide_d = nullptr; // a switch workspace per wave (switches of KS + 1 to KW ports)
    vector<int> queue_h, launched;
    vector<const rd::Sim *> sims_h;
    rd::LaunchCfg L = {};
};
constexpr int RSRV_SMALL_ITEMS = 32; // worlds of at most this many work items are small (teams of up to 8)
constexpr int RSRV_CAP = 4096;       // requests one launch takes at most
struct trd_server
{
    trd_server_cfg cfg = {};
    Arena arena;
    void *arena_mem = nullptr;
    hipStream_t up = nullptr;
    RSrvLane lane[2]; // small, large
    int max_blocks = 0, hpw = 16;
    uint64_t launches = 0;
    vector<RSrvReq> reqs;
    vector<int> waiting;
    size_t wait_head = 0;
    vector<int> redo; // requests to run again with larger tail rings (admitted before the waiting ones)
    ~trd_server()
    {
        for (auto &l : lane)
            if (l.ks)
                (void)hipStreamSynchronize(l.ks);
        for (auto &r : reqs)
            delete r.b;
        for (auto &l : lane)
        {
            for (void *p : {(void *)l.queue_d, (void *)l.qnext_d, (void *)l.teams_d, (void *)l.cfg_d, (void *)l.sims_d, (void *)l.wide_d})
                if (p)
                    (void)hipFree(p);
            if (l.done_ev)
                (void)hipEventDestroy(l.done_ev);
            if (l.ks)
                (void)hipStreamDestroy(l.ks);
        }
        if (arena_mem)
            (void)hipFree(arena_mem);
        if (up)
            (void)hipStreamDestroy(up);
    }
};

extern "C" int trd_server_create(const trd_server_cfg *cfg, trd_server **out)
{
    if (!out)
        return fail(TW_EINVAL, "trd_server_create: null output");
    int ndev = 0, dev = 0, cus = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(TW_EDEVICE, "no GPU: the RDMA model runs on the GPU only");
    std::unique_ptr<trd_server> s(new trd_server);
    if (cfg)
        s->cfg = *cfg;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return fail(TW_EDEVICE, "hipDeviceGetAttribute failed");
    const int resident = rd::launch_capacity(256);
    if (resident <= 0)
        return fail(TW_EDEVICE, "the rdma kernel fits no workgroup on the GPU");
    s->max_blocks = s->cfg.max_blocks > 0 ? std::min(s->cfg.max_blocks, resident) : std::max(2, std::min(resident, cus / 8));
    if (s->max_blocks < 2)
        return fail(TW_EINVAL, "max_blocks: at least 2 (one per class of worlds)");
    s->hpw = s->cfg.hosts_per_wave > 0 ? std::min(s->cfg.hosts_per_wave, 64) : 16;
    s->lane[0].blocks = std::max(1, s->max_blocks / 4);
    s->lane[1].blocks = s->max_blocks - s->lane[0].blocks;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess)
        return fail(TW_EDEVICE, "hipMemGetInfo failed");
    const size_t want = s->cfg.arena_bytes ? s->cfg.arena_bytes : fr / 2;
    if (hipMalloc(&s->arena_mem, want) != hipSuccess)
        return fail(TW_ENOMEM, "arena of %zu bytes", want);
    s->arena.base = (char *)s->arena_mem;
    s->arena.size = want;
    s->arena.free_[0] = want;
    if (hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess)
        return fail(TW_EDEVICE, "hipStreamCreate failed");
    for (auto &l : s->lane)
    {
        if (hipStreamCreateWithFlags(&l.ks, hipStreamNonBlocking) != hipSuccess || hipEventCreateWithFlags(&l.done_ev, hipEventDisableTiming) != hipSuccess)
            return fail(TW_EDEVICE, "hipStreamCreate failed");
        if (hipMalloc((void **)&l.queue_d, sizeof(int) * RSRV_CAP) != hipSuccess || hipMalloc((void **)&l.qnext_d, sizeof(unsigned)) != hipSuccess ||
            hipMalloc((void **)&l.teams_d, sizeof(rd::TeamState) * l.blocks) != hipSuccess ||
            hipMalloc((void **)&l.cfg_d, sizeof(rd::LaunchCfg)) != hipSuccess || hipMalloc((void **)&l.sims_d, sizeof(void *) * RSRV_CAP) != hipSuccess ||
            hipMalloc((void **)&l.wide_d, sizeof(rd::SwWorkWide) * (size_t)l.blocks * rd::LAUNCH_WAVES) != hipSuccess)
            return fail(TW_ENOMEM, "hipMalloc failed");
    }
    *out = s.release();
    return TW_OK;
}

extern "C" int trd_server_submit(trd_server *s, const trd_world *w, int32_t *id)
{
    if (!s || !w || !id)
        return fail(TW_EINVAL, "trd_server_submit: null argument");
    RSrvReq r;
    r.w.reset(new trd_world(*w));
    r.items = (w->H + s->hpw - 1) / s->hpw + w->S;
    s->reqs.push_back(std::move(r));
    *id = (int32_t)s->reqs.size() - 1;
    s->waiting.push_back(*id);
    return TW_OK;
}

// the end of a class's launch: its requests' worlds that finished
static int rsrv_collect(trd_server &s, RSrvLane &l)
{
    vector<rd::WorldRun> runs(l.launched.size());
    for (size_t i = 0; i < l.launched.size(); i++)
        if (hipMemcpyAsync(&runs[i], s.reqs[l.launched[i]].b->s.run, sizeof(rd::WorldRun), hipMemcpyDeviceToHost, l.ks) != hipSuccess)
            return fail(TW_EDEVICE, "hipMemcpy failed");
    // (each request's lost-tail counter: one world per request batch)
    vector<uint64_t> lost(l.launched.size(), 0);
    for (size_t i = 0; i < l.launched.size(); i++)
        if (hipMemcpyAsync(&lost[i], s.reqs[l.launched[i]].b->s.ev + rd::EV_LOST, sizeof(uint64_t), hipMemcpyDeviceToHost, l.ks) !=
            hipSuccess)
            return fail(TW_EDEVICE, "hipMemcpy failed");
    vector<rd::TeamState> ts(l.blocks);
    if (hipMemcpyAsync(ts.data(), l.teams_d, sizeof(rd::TeamState) * l.blocks, hipMemcpyDeviceToHost, l.ks) != hipSuccess ||
        hipStreamSynchronize(l.ks) != hipSuccess)
        return fail(TW_EDEVICE, "the rdma kernel failed: %s", hipGetErrorString(hipGetLastError()));
    bool err = false;
    for (auto &t : ts)
        err |= t.err != 0;
    for (size_t i = 0; i < l.launched.size(); i++)
    {
        RSrvReq &r = s.reqs[l.launched[i]];
        if (err)
            r.state = TW_REQ_FAILED;
        else if (runs[i].finished && lost[i] && r.tail_lg < RSRV_TAIL_MAX)
        {
            // a tie met an overwritten tail node: the request runs again from the start with larger
            // rings (2^24, then 2^26 nodes), as the reference reruns itself (src/rdma/rmain.cu:2332-2363)
            delete r.b; // (this class's launch has ended; no kernel uses it)
            r.b = nullptr;
            r.tail_lg = r.tail_lg < 24 ? 24 : RSRV_TAIL_MAX;
            r.state = TW_REQ_WAITING;
            s.redo.push_back(l.launched[i]);
        }
        else if (runs[i].finished)
            r.state = TW_REQ_DONE;
    }
    l.launched.clear();
    return err ? fail(TW_EDEVICE, "a team barrier timed out") : TW_OK;
}

// a launch of class c over its resident, unfinished requests
static int rsrv_launch(trd_server &s, int c)
{
    RSrvLane &l = s.lane[c];
    l.queue_h.clear();
    l.sims_h.clear();
    l.launched.clear();
    int mx = 1;
    for (size_t i = 0; i < s.reqs.size() && (int)l.launched.size() < RSRV_CAP; i++)
    {
        RSrvReq &r = s.reqs[i];
        if (r.state != TW_REQ_RESIDENT || (r.items > RSRV_SMALL_ITEMS) != (c == 1))
            continue;
        l.queue_h.push_back((int)l.sims_h.size() << rd::QSHIFT);
        l.sims_h.push_back(r.b->sim_d);
        l.launched.push_back((int)i);
        mx = std::max(mx, r.items);
    }
    if (l.launched.empty())
        return TW_OK;
    rd::LaunchCfg &L = l.L;
    L = rd::LaunchCfg{};
    // (small worlds: a workgroup each, as many at once as the class's share holds; large: teams of up
    // to 16, so that several run at once)
    L.T = c == 0 ? 1 : std::max(1, std::min(std::min(16, (mx + 3) / 4), l.blocks));
    L.nq = (int)l.queue_h.size();
    L.hpw = s.hpw;
    L.queue = l.queue_d;
    L.qnext = l.qnext_d;
    L.teams = l.teams_d;
    L.sims = l.sims_d;
    L.wide = l.wide_d;
    const int nteams = std::max(1, std::min(L.nq, l.blocks / L.T));
    if (hipMemcpyAsync(l.queue_d, l.queue_h.data(), sizeof(int) * L.nq, hipMemcpyHostToDevice, l.ks) != hipSuccess ||
        hipMemcpyAsync(l.sims_d, l.sims_h.data(), sizeof(void *) * L.nq, hipMemcpyHostToDevice, l.ks) != hipSuccess ||
        hipMemcpyAsync(l.cfg_d, &L, sizeof L, hipMemcpyHostToDevice, l.ks) != hipSuccess ||
        hipMemsetAsync(l.qnext_d, 0, sizeof(unsigned), l.ks) != hipSuccess ||
        hipMemsetAsync(l.teams_d, 0, sizeof(rd::TeamState) * l.blocks, l.ks) != hipSuccess)
        return fail(TW_EDEVICE, "launch set-up failed");
    const int e = rd::launch_worlds(l.cfg_d, nteams * L.T, 256, l.ks);
    if (e)
        return fail(TW_EDEVICE, "rdma kernel launch failed (%d)", e);
    if (hipEventRecord(l.done_ev, l.ks) != hipSuccess)
        return fail(TW_EDEVICE, "hipEventRecord failed");
    l.in_flight = true;
    s.launches++;
    return TW_OK;
}

extern "C" int trd_server_poll(trd_server *s, trd_server_stats *st)
{
    if (!s)
        return fail(TW_EINVAL, "trd_server_poll: null server");
    int rc = TW_OK;
    for (auto &l : s->lane)
        if (l.in_flight)
        {
            const hipError_t q = hipEventQuery(l.done_ev);
            if (q == hipSuccess)
            {
                l.in_flight = false;
                if (int e = rsrv_collect(*s, l))
                    rc = e;
            }
            else if (q != hipErrorNotReady)
                return fail(TW_EDEVICE, "the rdma kernel failed: %s", hipGetErrorString(q));
        }
    // admissions: reruns first, then in arrival order within each size class while the arena has
    // room (a large request waiting for memory does not hold up small ones)
    bool blocked[2] = {false, false};
    size_t scan = s->wait_head;
    while (!s->redo.empty() || scan < s->waiting.size())
    {
        const bool rerun = !s->redo.empty();
        RSrvReq &r = s->reqs[rerun ? s->redo.front() : s->waiting[scan]];
        const int cls = r.items > RSRV_SMALL_ITEMS;
        if (!rerun && (r.state != TW_REQ_WAITING || blocked[cls]))
        {
            scan++;
            continue;
        }
        std::unique_ptr<trd_batch> b(new trd_batch);
        b->arena = &s->arena;
        b->st = s->up;
        b->own_stream = false;
        trd_pack_cfg pc = {};
        pc.q3cap = s->cfg.q3cap;
        pc.tail_log2 = r.tail_lg ? r.tail_lg : s->cfg.tail_log2 > 0 ? s->cfg.tail_log2 : 22;
        pc.tail_exact = r.tail_lg ? 1 : 0;
        trd_world *wp = r.w.get();
        const int e = pack(*b, &wp, 1, pc);
        if (e == TW_ENOMEM)
        {
            (void)hipStreamSynchronize(s->up); // (the partial batch's memsets end before its memory returns)
            b.reset();
            if (s->arena.used > 0)
            {
                // it waits for memory (a rerun holds up every admission, as it did before)
                if (rerun)
                    break;
                blocked[cls] = true;
                scan++;
                continue;
            }
            // (it does not fit even an empty arena: it never will)
        }
        if (rerun)
            s->redo.erase(s->redo.begin());
        else
            scan++;
        if (e)
        {
            r.state = TW_REQ_FAILED;
            continue;
        }
        r.b = b.release();
        r.state = TW_REQ_RESIDENT;
    }
    while (s->wait_head < s->waiting.size() && s->reqs[s->waiting[s->wait_head]].state != TW_REQ_WAITING)
        s->wait_head++;
    for (int c = 0; c < 2; c++)
        if (!s->lane[c].in_flight)
            if (int e = rsrv_launch(*s, c))
                return e;
    if (st)
    {
        *st = trd_server_stats{};
        for (auto &r : s->reqs)
        {
            st->waiting += r.state == TW_REQ_WAITING;
            st->resident += r.state == TW_REQ_RESIDENT;
            st->done += r.state == TW_REQ_DONE || r.state == TW_REQ_TAKEN;
            st->failed += r.state == TW_REQ_FAILED;
        }
        st->arena_used = s->arena.used;
        st->arena_bytes = s->arena.size;
        st->launches = s->launches;
    }
    return rc;
}

extern "C" int trd_server_state(const trd_server *s, int32_t id, int32_t *state)
{
    if (!s || id < 0 || id >= (int32_t)s->reqs.size() || !state)
        return fail(TW_EINVAL, "trd_server_state: bad request %d", id);
    *state = s->reqs[id].state;
    return TW_OK;
}

extern "C" int trd_server_result(trd_server *s, int32_t id, int64_t *start_ns, int64_t *fct_ns, uint64_t *ev40)
{
    if (!s || id < 0 || id >= (int32_t)s->reqs.size())
        return fail(TW_EINVAL, "trd_server_result: bad request %d", id);
    RSrvReq &r = s->reqs[id];
    if (r.state == TW_REQ_FAILED)
    {
        // a failed request's memory returns to the arena (its class's launch has ended: a request
        // is marked failed only when the launch's end is collected, or before it is admitted)
        delete r.b;
        r.b = nullptr;
        r.w.reset();
        r.state = TW_REQ_TAKEN;
        return fail(TW_EDEVICE, "trd_server_result: request %d failed (a team barrier timed out, or its world could not be packed)", id);
    }
    if (r.state != TW_REQ_DONE)
        return fail(TW_ESTATE, "trd_server_result: request %d is not done (state %d)", id, r.state);
    int e = trd_fct_get(r.b, 0, start_ns, fct_ns);
    if (!e && ev40)
        e = trd_counters_get(r.b, ev40);
    if (e)
        return e;
    delete r.b; // (its memory returns to the arena; no kernel uses it any more)
    r.b = nullptr;
    r.w.reset();
    r.state = TW_REQ_TAKEN;
    return TW_OK;
}

extern "C" void trd_server_destroy(trd_server *s) { delete s; }
