// TEST INFRASTRUCTURE: the RDMA serving loop (trd_server_*) with the slot kernel's steps run on the
// CPU (rdma_cpu.cpp). Submits reference jobs to a server whose arena holds only a few of them at
// once (admission waits for memory), polls until every request is done, and writes each request's
// flow completions as "id start_ns fct_ns" lines to OUTDIR/NAME.fct for comparison with the
// reference's. usage: server_cpu CONFIG INPUTDIR OUTDIR ARENA_MB NAME TOPO FLOWS KV [NAME TOPO FLOWS KV ...]
// (KV: comma-separated KEY=VALUE overrides, or -)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../../include/twins.h"
#include "../../../include/twins_rdma.h"

int main(int argc, char **argv)
{
    if (argc < 9 || (argc - 5) % 4)
        return fprintf(stderr, "usage: server_cpu CONFIG INPUTDIR OUTDIR ARENA_MB NAME TOPO FLOWS KV ...\n"), 1;
    const std::string cfg = argv[1], in = argv[2], out = argv[3];
    trd_server_cfg sc = {};
    sc.arena_bytes = (uint64_t)atoll(argv[4]) << 20;
    sc.tail_log2 = 16;
    trd_server *s = nullptr;
    if (trd_server_create(&sc, &s))
        return fprintf(stderr, "%s\n", trd_last_error()), 1;
    std::vector<std::string> names;
    std::vector<int> flows;
    for (int i = 5; i < argc; i += 4)
    {
        std::vector<std::string> kv;
        std::string k = argv[i + 3];
        for (size_t b = 0, c; k != "-" && b <= k.size(); b = c + 1)
        {
            c = k.find(',', b);
            if (c == std::string::npos)
                c = k.size();
            kv.push_back(k.substr(b, c - b));
        }
        std::vector<const char *> ov;
        for (auto &x : kv)
            ov.push_back(x.c_str());
        trd_world *w = nullptr;
        if (trd_world_create(cfg.c_str(), ov.data(), (int)ov.size(), (in + "/" + argv[i + 1]).c_str(), (in + "/" + argv[i + 2]).c_str(), 0, 0, &w))
            return fprintf(stderr, "%s\n", trd_last_error()), 1;
        int32_t id, f;
        trd_world_info(w, nullptr, nullptr, &f, nullptr, nullptr);
        if (trd_server_submit(s, w, &id))
            return fprintf(stderr, "%s\n", trd_last_error()), 1;
        trd_world_destroy(w); // (the server keeps a copy)
        names.push_back(argv[i]);
        flows.push_back(f);
    }
    int max_waiting = 0, polls = 0;
    for (;;)
    {
        trd_server_stats st;
        if (trd_server_poll(s, &st))
            return fprintf(stderr, "%s\n", trd_last_error()), 1;
        polls++;
        max_waiting = st.waiting > max_waiting ? st.waiting : max_waiting;
        for (int id = 0; id < (int)names.size(); id++)
        {
            int32_t state;
            trd_server_state(s, id, &state);
            if (state == TW_REQ_FAILED)
                return fprintf(stderr, "request %d failed\n", id), 1;
            if (state != TW_REQ_DONE)
                continue;
            std::vector<int64_t> a(flows[id]), b(flows[id]);
            uint64_t ev[40];
            if (trd_server_result(s, id, a.data(), b.data(), ev))
                return fprintf(stderr, "%s\n", trd_last_error()), 1;
            FILE *fo = fopen((out + "/" + names[id] + ".fct").c_str(), "w");
            for (int i = 0; i < flows[id]; i++)
                if (b[i] >= 0)
                    fprintf(fo, "%d %lld %lld\n", i, (long long)a[i], (long long)b[i]);
            fclose(fo);
        }
        if (st.done + st.failed == (int)names.size())
        {
            printf("requests %d done, launches %llu, polls %d, most waiting %d, arena used %llu\n", st.done, (unsigned long long)st.launches,
                   polls, max_waiting, (unsigned long long)st.arena_used);
            break;
        }
    }
    trd_server_destroy(s);
    return 0;
}
