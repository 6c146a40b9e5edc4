// twins-rdma: the RDMA model's command line over libtwins_core (include/twins_rdma.h), with the
// reference's arguments, output files and summary lines (src/rdma/rmain.cu:1-42, 1023-1254,
// 2178-2258, 2316-2373):
//
//   twins-rdma CONFIG --topo=F --flows=F --out=F [--rate=100Gbps --delay=1000] [--q3cap=N]
//              [--tail=log2 | --tail-auto=log2] [--txtrace=F --from=ns --until=ns] [--sent=F] [KEY=VALUE ...]
//   twins-rdma CONFIG --batch=JOBFILE --outdir=DIR [--path=DIR:DIR...] [options] [KEY=VALUE ...]
//
// out: "id start_ns fct_ns" per finished flow, by id. Exit status: 0 every flow done, 2 some
// unfinished at the stop time, 3 an overflow or a late event (the results are not exact), 1 an
// argument or input error. (The reference's launch-shape flags --serial, --spread, --win, --gwin,
// --persist, --rank* and --dcap are accepted and change nothing here.)
//
// Chain tails (src/rdma/rmain.cu:25-31, 2326-2363): each world keeps the older runs of its event
// chains in a ring of tail nodes, 2^(log2 entities + 14) nodes up to 2^24 unless --tail=N sets it
// to 2^N. A tie that meets an overwritten node is counted "on lost tails"; the run is then made
// again in this process, on a new batch whose rings hold 2^26 nodes, then 2^27 (--tail-auto=N: start
// at 2^N and still grow; --tail=N: never grow), as the reference re-executes itself.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/twins.h"
#include "../../include/twins_rdma.h"

using std::string;
using std::vector;

constexpr int TAIL_MAX = 27; // the largest ring a rerun asks for (2^27 nodes, the reference's)

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

struct Job
{
    string name, topo, flows;
    vector<string> kv;
};

int main(int argc, char **argv)
{
    if (argc < 2)
    {
        fprintf(stderr, "usage: twins-rdma CONFIG --topo=F --flows=F --out=F [KEY=VALUE ...]\n"
                        "       twins-rdma CONFIG --batch=JOBFILE --outdir=DIR [KEY=VALUE ...]\n");
        return 1;
    }
    const string config = argv[1];
    string topo, flows, out = "fct.txt", batch, outdir = ".", dirs, rate_s = "100Gbps";
    int64_t delay = 1000, tx_from = 0, tx_until = INT64_MAX;
    string txtrace, sent;
    trd_pack_cfg pc = {};
    bool tail_set = false;
    vector<string> kv;
    for (int i = 2; i < argc; i++)
    {
        const string a = argv[i];
        const size_t eq = a.find('=');
        const string k = a.substr(0, eq), v = eq == string::npos ? "" : a.substr(eq + 1);
        if (k == "--topo")
            topo = v;
        else if (k == "--flows")
            flows = v;
        else if (k == "--out")
            out = v;
        else if (k == "--batch")
            batch = v;
        else if (k == "--outdir")
            outdir = v;
        else if (k == "--path")
            dirs = v;
        else if (k == "--rate")
            rate_s = v;
        else if (k == "--delay")
            delay = atoll(v.c_str());
        else if (k == "--q3cap")
            pc.q3cap = atoi(v.c_str());
        else if (k == "--tail" || k == "--tail-auto")
        {
            pc.tail_log2 = atoi(v.c_str());
            pc.tail_exact = 1;
            tail_set = k == "--tail";
            if (pc.tail_log2 < 4 || pc.tail_log2 > 30)
            {
                fprintf(stderr, "%s: log2 of the tail ring from 4 to 30\n", k.c_str());
                return 1;
            }
        }
        else if (k == "--txtrace")
            txtrace = v;
        else if (k == "--from")
            tx_from = atoll(v.c_str());
        else if (k == "--until")
            tx_until = atoll(v.c_str());
        else if (k == "--sent")
            sent = v;
        else if (k == "--astra" || k.compare(0, 5, "--nvl") == 0 || k.compare(0, 5, "--isl") == 0)
        {
            fprintf(stderr, "%s: not modelled in this build\n", k.c_str());
            return 1;
        }
        else if (k == "--serial")
            pc.serial = 1; // every switch slot in order on one lane (src/rdma/rmain.cu:1090-1091, 1949)
        else if (k == "--team_blocks") // (launch shape, no counterpart in the reference: workgroups per world)
            pc.team_blocks = atoi(v.c_str());
        else if (k == "--hosts_per_wave") // (likewise: hosts side by side on a wave's lanes)
            pc.hosts_per_wave = atoi(v.c_str());
        else if (k == "--spread" || k == "--win" || k == "--gwin" || k == "--persist" || k.compare(0, 6, "--rank") == 0 ||
                 k == "--dcap") // (--dcap sizes DAG message buffers; DAG flow files are refused)
            continue;
        else if (k.compare(0, 2, "--") != 0 && eq != string::npos)
            kv.push_back(a);
        else
        {
            fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 1;
        }
    }
    const uint64_t bps = parse_rate(rate_s);
    auto find_file = [&](const string &f) {
        size_t b = 0;
        while (!dirs.empty() && b <= dirs.size())
        {
            size_t c = dirs.find(':', b);
            const string p = dirs.substr(b, c == string::npos ? string::npos : c - b) + "/" + f;
            if (std::ifstream(p))
                return p;
            if (c == string::npos)
                break;
            b = c + 1;
        }
        return f;
    };
    if (!txtrace.empty() && !batch.empty())
    {
        fprintf(stderr, "--txtrace traces one world (no --batch)\n"); // (src/rdma/rmain.cu:1199-1202)
        return 1;
    }
    vector<Job> jobs;
    if (batch.empty())
        jobs.push_back({"run", topo, flows, kv});
    else
    {
        std::ifstream f(batch);
        if (!f)
        {
            fprintf(stderr, "cannot open %s\n", batch.c_str());
            return 1;
        }
        string line;
        while (std::getline(f, line))
        {
            std::istringstream is(line);
            Job j;
            if (!(is >> j.name) || j.name[0] == '#' || !(is >> j.topo >> j.flows))
                continue;
            j.kv = kv;
            string x;
            while (is >> x)
            {
                if (x.find('=') == string::npos)
                {
                    fprintf(stderr, "%s: bad override %s\n", j.name.c_str(), x.c_str());
                    return 1;
                }
                j.kv.push_back(x);
            }
            j.topo = find_file(j.topo);
            j.flows = find_file(j.flows);
            jobs.push_back(j);
        }
        if (jobs.empty())
        {
            fprintf(stderr, "%s: no jobs\n", batch.c_str());
            return 1;
        }
    }
    vector<trd_world *> worlds;
    int F = 0, H = 0, S = 0;
    for (auto &j : jobs)
    {
        vector<const char *> ov;
        for (auto &x : j.kv)
            ov.push_back(x.c_str());
        trd_world *w = nullptr;
        if (trd_world_create(config.c_str(), ov.data(), (int)ov.size(), j.topo.c_str(), j.flows.c_str(), bps, delay, &w))
        {
            fprintf(stderr, "%s: %s\n", j.name.c_str(), trd_last_error());
            return 1;
        }
        int32_t h, s, f;
        uint64_t rtt;
        uint32_t bdp;
        trd_world_info(w, &h, &s, &f, &rtt, &bdp);
        if (jobs.size() == 1)
            printf("maxRtt=%llu maxBdp=%u\n", (unsigned long long)rtt, bdp);
        F += f, H += h, S += s;
        worlds.push_back(w);
    }
    if (jobs.size() > 1)
        printf("%d worlds: %d hosts, %d switches, %d flows\n", (int)jobs.size(), H, S, F);
    fflush(stdout);
    // a run of batch b: the output files and summary lines; the exit status (-1: an error)
    auto run_once = [&](trd_batch *b) -> int {
        if (!txtrace.empty() && trd_txtrace_enable(b, tx_from, tx_until, 0))
        {
            fprintf(stderr, "%s\n", trd_last_error());
            return -1;
        }
        trd_run_stats st;
        if (trd_batch_run(b, &st))
        {
            fprintf(stderr, "%s\n", trd_last_error());
            return -1;
        }
        uint64_t ev[40];
        trd_counters_get(b, ev);
        for (size_t wi = 0; wi < jobs.size(); wi++)
        {
            int32_t f = 0;
            trd_world_info(worlds[wi], nullptr, nullptr, &f, nullptr, nullptr);
            vector<int64_t> start(f), fct(f);
            trd_fct_get(b, (int)wi, start.data(), fct.data());
            const string path = batch.empty() ? out : outdir + "/" + jobs[wi].name + ".fct";
            FILE *fo = fopen(path.c_str(), "w");
            if (!fo)
            {
                fprintf(stderr, "cannot write %s\n", path.c_str());
                return -1;
            }
            for (int i = 0; i < f; i++)
                if (fct[i] >= 0)
                    fprintf(fo, "%d %lld %lld\n", i, (long long)start[i], (long long)fct[i]);
            fclose(fo);
        }
        printf("flows done %u of %d; simulated %.6f s in %.3f s wall (%d launches)\n", st.flows_done, F, st.sim_ns * 1e-9, st.wall_s,
               st.launches);
        printf("drops %llu marks %llu pauses %llu resumes %llu nacks %llu | undecided ties %llu, on lost tails %llu | overflows: tx %llu queue "
               "%llu ackq %llu qps %llu due %llu | acks for done QPs %llu\n",
               (unsigned long long)ev[0], (unsigned long long)ev[1], (unsigned long long)ev[2], (unsigned long long)ev[3],
               (unsigned long long)ev[11], (unsigned long long)ev[8], (unsigned long long)ev[9], (unsigned long long)ev[4],
               (unsigned long long)ev[5], (unsigned long long)ev[6], (unsigned long long)ev[7], (unsigned long long)ev[12],
               (unsigned long long)ev[10]);
        if ((ev[33] || ev[34]) && st.slots > 0)
        {
            // (profiling builds: wave-time of host items and of switch items summed over the team, and
            // the first workgroup's wait at the barrier, per slot; constant clock of 100 MHz)
            const double k = 0.01 / (double)st.slots;
            printf("prof: per slot: host items %.2f wave-us, switch items %.2f wave-us, barrier %.2f us (%lld slots)\n",
                   ev[33] * k, ev[34] * k, ev[35] * k, (long long)st.slots);
            // the switch slots' phases (ports in, routes, events, ports out), their arrivals, and in the
            // events of safe slots the egress merge, TransmitCompletes and arrivals (a wave's time each)
            printf("prof: switch phases per slot (wave-us): in %.2f, routes %.2f, events %.2f, out %.2f; arrivals %.1f; "
                   "merge %.2f, tx-complete %.2f, arrival events %.2f; safe slots %llu, serial %llu\n",
                   ev[24] * k, ev[25] * k, ev[26] * k, ev[27] * k, ev[28] / (double)st.slots, ev[29] * k, ev[30] * k, ev[31] * k,
                   (unsigned long long)ev[36], (unsigned long long)ev[37]);
        }
        if (!txtrace.empty())
        {
            uint64_t n = 0, seen = 0;
            if (trd_txtrace_write(b, 0, txtrace.c_str(), &n, &seen))
            {
                fprintf(stderr, "%s\n", trd_last_error());
                return -1;
            }
            if (seen > n)
                fprintf(stderr, "txtrace: %llu records, %llu kept\n", (unsigned long long)seen, (unsigned long long)n);
        }
        if (!sent.empty())
        {
            // "id ns" per flow at its first SendComplete: the reference records it only under the SimAI
            // transmit rule (src/rdma/rstep.cuh:1062-1069), which this build's worlds never use (SimAI
            // topologies are refused), so the file is empty as the reference's is for these worlds
            // (src/rdma/rmain.cu:2202-2219; with --batch, OUTDIR/JOB.sent per world)
            for (size_t wi = 0; wi < jobs.size(); wi++)
            {
                const string path = batch.empty() ? sent : outdir + "/" + jobs[wi].name + ".sent";
                FILE *fs = fopen(path.c_str(), "w");
                if (!fs)
                {
                    fprintf(stderr, "cannot write %s\n", path.c_str());
                    return -1;
                }
                fclose(fs);
            }
        }
        const bool bad = ev[4] || ev[5] || ev[6] || ev[7] || ev[12] || ev[32];
        if (bad)
            fprintf(stderr, "warning: capacity overflow, results are not exact\n");
        if (ev[8] || ev[9])
            fprintf(stderr, "warning: ties the chains did not decide exactly (%llu undecided, %llu on lost tails; a larger --tail helps)\n",
                    (unsigned long long)ev[8], (unsigned long long)ev[9]);
        return bad ? 3 : st.flows_done == (uint32_t)F ? 0 : 2;
    };
    // one run; on lost tails (and no --tail) the same worlds again on a batch with larger rings
    for (int prev = -1;;)
    {
        trd_batch *b = nullptr;
        if (trd_batch_create(worlds.data(), (int)worlds.size(), &pc, &b))
        {
            fprintf(stderr, "%s\n", trd_last_error());
            // (a rerun whose larger rings do not fit leaves the last run's results in place)
            return prev >= 0 ? prev : 1;
        }
        const int rc = run_once(b);
        prev = rc;
        int32_t lg = 0;
        uint64_t lost = 0;
        trd_batch_tail_log2(b, &lg);
        {
            uint64_t ev[40];
            trd_counters_get(b, ev);
            lost = ev[9];
        }
        trd_batch_destroy(b);
        if (rc < 0)
            return 1;
        if (!lost || tail_set || lg >= TAIL_MAX)
        {
            for (auto *w : worlds)
                trd_world_destroy(w);
            return rc;
        }
        pc.tail_log2 = lg < 26 ? 26 : TAIL_MAX;
        pc.tail_exact = 1;
        fprintf(stderr, "lost tails: running again with --tail-auto=%d\n", pc.tail_log2);
        fflush(stdout);
        fflush(stderr);
        if (jobs.size() == 1)
        {
            // (the reference's re-executed run prints its set-up line again)
            int32_t h, sw, f;
            uint64_t rtt;
            uint32_t bdp;
            trd_world_info(worlds[0], &h, &sw, &f, &rtt, &bdp);
            printf("maxRtt=%llu maxBdp=%u\n", (unsigned long long)rtt, bdp);
        }
        else
            printf("%d worlds: %d hosts, %d switches, %d flows\n", (int)jobs.size(), H, S, F);
    }
}
