// `twins`: the drop-in command line of the TCP simulator (reference: src/main.cu). Same flags,
// worlds files, stdout summary lines and fct.csv output; the simulation runs through the C ABI of
// libtwins_core on the GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/twins.h"

static int usage()
{
    fprintf(stderr, "usage: twins (--k=<4..32> | --topo=leafspine --leaves=L --spines=S --hpl=N | --topo=graph --topofile=F)\n"
                    "         (--flows=<csv> [--out=csv] [--tcp=dctcp|newreno|cubic] [--ecnK=30] [--queue=200] [--fin=0|1]\n"
                    "          [--ecmp_from=csv [--hash_check=1]] [--dump_routes=file] [--spray=1 [--RngSeed=1 --RngRun=1]] [--chain_runs=9|25]\n"
                    "          [--qtrace=file --qtrace_from=ps --qtrace_until=ps] (every enqueue in [from, until))\n"
                    "          [--block_threads=256 --blocks=N --team_blocks=T --hosts_per_wave=H] (launch shape; 0: automatic)\n"
                    "          [--noalloc=1] (run with the library's static scratch, allocating nothing: at most 33 workgroups)\n"
                    "          | --worlds=<file>: one world per line, flows=..,tcp=..,ecnK=..,queue=..,fin=..,ecmp_from=..,out=..)\n"
                    "         leaf-spine: max(N+S, L) <= 32\n");
    return 1;
}

int main(int argc, char **argv)
{
    std::vector<std::string> common; // topology and defaults, passed to every world
    std::string flows, out = "fct.csv", worlds_file, tcp = "dctcp", ecmp_from;
    int runs = 9, threads = 0, blocks = 0, team = 0, hpw = 0;
    std::string ecn_k, queue, fin, qtrace;
    long long qtrace_from = 0, qtrace_until = 0;
    int hash_check = 0;
    int noalloc = 0; // run through tw_batch_run_noalloc (the library's static scratch)
    std::string dump_routes;
    for (int i = 1; i < argc; i++)
    {
        std::string a = argv[i];
        auto is = [&](const char *p) { return a.rfind(p, 0) == 0; };
        auto val = [&](const char *p) { return a.substr(strlen(p)); };
        if (is("--flows="))
            flows = val("--flows=");
        else if (is("--out="))
            out = val("--out=");
        else if (is("--worlds="))
            worlds_file = val("--worlds=");
        else if (is("--tcp="))
            tcp = val("--tcp=");
        else if (is("--ecmp_from="))
            ecmp_from = val("--ecmp_from=");
        else if (is("--ecnK="))
            ecn_k = val("--ecnK=");
        else if (is("--queue="))
            queue = val("--queue=");
        else if (is("--fin="))
            fin = val("--fin=");
        else if (is("--chain_runs="))
            runs = atoi(val("--chain_runs=").c_str());
        else if (is("--block_threads="))
            threads = atoi(val("--block_threads=").c_str());
        else if (is("--blocks="))
            blocks = atoi(val("--blocks=").c_str());
        else if (is("--team_blocks="))
            team = atoi(val("--team_blocks=").c_str());
        else if (is("--hosts_per_wave="))
            hpw = atoi(val("--hosts_per_wave=").c_str());
        else if (is("--qtrace="))
            qtrace = val("--qtrace=");
        else if (is("--qtrace_from="))
            qtrace_from = atoll(val("--qtrace_from=").c_str());
        else if (is("--qtrace_until="))
            qtrace_until = atoll(val("--qtrace_until=").c_str());
        else if (is("--dump_routes="))
            dump_routes = val("--dump_routes=");
        else if (is("--hash_check="))
            hash_check = atoi(val("--hash_check=").c_str());
        else if (is("--noalloc="))
            noalloc = atoi(val("--noalloc=").c_str());
        else if (a == "--general" || is("--graph_k=") || is("--host_lanes="))
            ; // the reference's kernel-shape and debugging switches: nothing to choose here
        else if (is("--k=") || is("--topo=") || is("--topofile=") || is("--fail=") || is("--leaves=") || is("--spines=") ||
                 is("--hpl=") || is("--spray=") || is("--RngSeed=") || is("--RngRun=") || is("--max_slots=") || is("--rtt_ring=") ||
                 is("--ooo_segs="))
            common.push_back(a);
        else
        {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 1;
        }
    }
    if ((flows.empty() && worlds_file.empty()) || (tcp != "dctcp" && tcp != "newreno" && tcp != "cubic"))
        return usage();
    struct World
    {
        std::vector<std::string> args;
        std::string out;
    };
    std::vector<World> worlds;
    auto base = [&]() {
        std::vector<std::string> v = common;
        v.push_back("--tcp=" + tcp);
        if (!ecn_k.empty())
            v.push_back("--ecnK=" + ecn_k);
        if (!queue.empty())
            v.push_back("--queue=" + queue);
        if (!fin.empty())
            v.push_back("--fin=" + fin);
        return v;
    };
    if (worlds_file.empty())
    {
        World w{base(), out};
        w.args.push_back("--flows=" + flows);
        if (!ecmp_from.empty())
            w.args.push_back("--ecmp_from=" + ecmp_from);
        worlds.push_back(w);
    }
    else
    {
        // one world per line: key=value,... (src/main.cu:277-325)
        std::ifstream in(worlds_file);
        if (!in)
        {
            fprintf(stderr, "cannot open %s\n", worlds_file.c_str());
            return 1;
        }
        std::string line;
        while (std::getline(in, line))
        {
            if (line.empty() || line[0] == '#')
                continue;
            World w{base(), ""};
            bool has_flows = false;
            std::stringstream ss(line);
            for (std::string kv; std::getline(ss, kv, ',');)
            {
                size_t e = kv.find('=');
                std::string key = kv.substr(0, e), v = e == std::string::npos ? "" : kv.substr(e + 1);
                if (key == "flows")
                    w.args.push_back("--flows=" + v), has_flows = true;
                else if (key == "tcp")
                {
                    if (v == "dctcp" || v == "newreno" || v == "cubic")
                        w.args.push_back("--tcp=" + v);
                }
                else if (key == "ecnK" || key == "queue" || key == "fin" || key == "ecmp_from")
                    w.args.push_back("--" + key + "=" + v);
                else if (key == "out")
                    w.out = v;
                else
                {
                    fprintf(stderr, "%s: bad entry %s\n", worlds_file.c_str(), kv.c_str());
                    return 1;
                }
            }
            if (!has_flows || w.out.empty())
            {
                fprintf(stderr, "%s: every world needs flows= and out=\n", worlds_file.c_str());
                return 1;
            }
            worlds.push_back(w);
        }
    }
    std::vector<tw_world *> hw;
    for (auto &w : worlds)
    {
        std::vector<const char *> av;
        for (auto &s : w.args)
            av.push_back(s.c_str());
        tw_world *x = nullptr;
        if (tw_world_create_args((int)av.size(), av.data(), &x))
        {
            fprintf(stderr, "%s\n", tw_last_error());
            return 1;
        }
        hw.push_back(x);
    }
    if (hw.empty())
    {
        fprintf(stderr, "%s: no worlds\n", worlds_file.c_str());
        return 1;
    }
    // --dump_routes=F: the first world's routes, written for a topology-file world only and
    // otherwise ignored, as the reference writes them in its graph branch (src/main.cu:466-486)
    bool graph = false;
    for (auto &a : worlds[0].args)
        graph |= a == "--topo=graph";
    if (!dump_routes.empty() && graph && tw_world_dump_routes(hw[0], dump_routes.c_str()))
    {
        fprintf(stderr, "%s\n", tw_last_error());
        return 1;
    }
    if (hash_check)
    {
        // --hash_check=1: the computed ECMP hashes against --ecmp_from's, then stop (src/main.cu:763-768)
        int same = 0, n = 0;
        for (tw_world *x : hw)
        {
            int32_t a, c;
            if (tw_world_hash_check(x, &a, &c))
            {
                fprintf(stderr, "%s\n", tw_last_error());
                return 1;
            }
            same += a, n += c;
        }
        printf("ECMP hashes: %d of %d connections as ns-3's\n", same, n);
        return same == n ? 0 : 1;
    }
    tw_pack_cfg cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.chain_runs = runs;
    cfg.block_threads = threads;
    cfg.blocks = blocks;
    cfg.team_blocks = team;
    cfg.hosts_per_wave = hpw;
    tw_batch *b = nullptr;
    if (tw_batch_create(hw.data(), (int)hw.size(), &cfg, &b))
    {
        fprintf(stderr, "%s\n", tw_last_error());
        return 1;
    }
    // the queue trace (src/main.cu:169-174, 985-994)
    if (!qtrace.empty() && tw_qtrace_enable(b, qtrace_from, qtrace_until, 0))
    {
        fprintf(stderr, "%s\n", tw_last_error());
        return 1;
    }
    tw_run_stats st;
    if ((noalloc ? tw_batch_run_noalloc : tw_batch_run)(b, &st))
    {
        fprintf(stderr, "%s\n", tw_last_error());
        return 1;
    }
    // the report (src/main.cu:1064-1075), summed over worlds
    unsigned long long ev[32] = {0};
    long long slots = 0;
    int M = 0;
    std::vector<int> undecided;
    for (size_t i = 0; i < hw.size(); i++)
    {
        uint64_t e[32];
        tw_counters_get(b, (int)i, e);
        for (int j = 0; j < 31; j++)
            ev[j] = j == 13 ? std::max<unsigned long long>(ev[j], e[j]) : ev[j] + e[j];
        slots = std::max<long long>(slots, (long long)e[31]);
        if (e[25])
            undecided.push_back((int)i);
        tw_world_info info;
        tw_world_info_get(hw[i], &info);
        M += info.flows;
    }
    printf("Simulation Duration: %.6fs\n", st.wall_s);
    printf("Executed Slots: %lld\n", slots);
    printf("flows %d completed %llu wall %.4f s drops %llu marks %llu retx %llu rto %llu | tx-cap %llu active-ovf %llu ooo-ovf %llu due-ovf %llu"
           " | retx by state open %llu disorder %llu cwr %llu recovery %llu loss %llu | max-inflight %llu segs rtt-ovf %llu ring-ovf %llu"
           " | undecided-ties %llu\n",
           M, (unsigned long long)st.flows_done, st.wall_s, ev[0], ev[1], ev[2], ev[3], ev[4], ev[5], ev[6], ev[7], ev[8], ev[9], ev[10],
           ev[11], ev[12], ev[13], ev[14], ev[19], ev[25]);
    printf("packet-events %llu kernel %.6f s launches %lld\n", (unsigned long long)st.packet_events, st.kernel_s, (long long)st.launches);
    if (ev[27]) // a TW_PROF build: wall-clock ticks (100 MHz) summed over slots
        printf("prof: per slot: period %.2f us, slowest workgroup: hosts %.2f us, switches %.2f us, work %.2f us\n", ev[27] / 100.0 / slots,
               ev[28] / 100.0 / slots, ev[29] / 100.0 / slots, ev[30] / 100.0 / slots);
    if (ev[27]) // switch items' phases, summed over items (not per slot maxima)
        printf("prof: switch phases (item-us per slot): window+spray %.2f, count %.2f, scatter %.2f, egress %.2f\n",
               ev[20] / 100.0 / slots, ev[21] / 100.0 / slots, ev[22] / 100.0 / slots, ev[23] / 100.0 / slots);
    if (ev[27]) // host lanes' phases, summed over hosts (not per slot maxima)
        printf("prof: host phases (lane-us per slot): set-up %.2f, loop %.2f (events %.2f), end %.2f; events per slot %.2f\n",
               ev[15] / 100.0 / slots, ev[16] / 100.0 / slots, ev[17] / 100.0 / slots, ev[18] / 100.0 / slots, (double)ev[24] / slots);
    if (ev[5] || ev[6] || ev[7] || ev[14] || ev[19])
        fprintf(stderr, "warning: a capacity overflowed (active/ooo/due/rtt/ring-ovf); the results may differ from ns-3\n");
    if (ev[25])
        fprintf(stderr, "warning: %llu ties between simultaneous events were undecided by their recorded histories (%d runs); "
                        "the results may differ from ns-3: rerun with --chain_runs=25 (longer histories)\n",
                ev[25], runs);
    if (!qtrace.empty())
    {
        uint64_t n = 0;
        if (tw_qtrace_write(b, qtrace.c_str(), &n))
        {
            fprintf(stderr, "%s\n", tw_last_error());
            return 1;
        }
        printf("qtrace: %llu records\n", (unsigned long long)n); // (src/main.cu:1077-1100)
    }
    // fct.csv per world (src/main.cu:1143-1153)
    for (size_t i = 0; i < hw.size(); i++)
    {
        tw_world_info info;
        tw_world_info_get(hw[i], &info);
        std::vector<int64_t> est(info.flows), done(info.flows);
        tw_fct_get(b, (int)i, est.data(), done.data());
        // the flows as read: re-read the file for src, dst, size, start
        std::string fl;
        for (auto &a : worlds[i].args)
            if (a.rfind("--flows=", 0) == 0)
                fl = a.substr(8);
        std::ifstream in(fl);
        FILE *o = fopen(worlds[i].out.c_str(), "w");
        if (!o)
        {
            fprintf(stderr, "cannot write %s\n", worlds[i].out.c_str());
            return 1;
        }
        fprintf(o, "src,dst,flow_size,start_ns,established_ns,completed_ns\n");
        std::string line;
        int j = 0;
        const int H = info.hosts;
        while (std::getline(in, line) && j < info.flows)
        {
            if (line.empty() || !isdigit((unsigned char)line[0]))
                continue;
            std::vector<long long> v;
            std::stringstream ss(line);
            for (std::string f; std::getline(ss, f, ',');)
                v.push_back(std::stoll(f));
            if (v.size() == 3)
                v = {v[0], (v[0] + H / 2) % H, v[1], v[2]};
            if (v.size() < 4)
                continue;
            fprintf(o, "%lld,%lld,%lld,%lld,%lld,%lld\n", v[0], v[1], v[2], v[3], (long long)est[j], (long long)done[j]);
            j++;
        }
        fclose(o);
    }
    if (hw.size() > 1)
    {
        printf("worlds %d\n", (int)hw.size());
        if (!undecided.empty())
        {
            printf("undecided-ties by world:");
            for (int w : undecided)
                printf(" %d", w);
            printf("\n");
        }
    }
    tw_batch_destroy(b);
    for (auto *w : hw)
        tw_world_destroy(w);
    return 0;
}
