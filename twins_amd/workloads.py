"""Synthetic inputs of the benchmark configurations, in the `twins` file formats.

The reference's generator (workloads/gen.py) has no lognormal or Pareto sizes, so these are
written here, with numpy's PCG64 seeded per configuration. Formats: flow CSV rows
`src,dst,flow_size,start_ns[,conn]` (a conn id < 0 is a connection of its own) and topology files
`hosts H / switches S / link A B` in the link order of the reference's FatTree generator (host
links, then aggregation-edge, then core-aggregation; workloads/topo.py:24-41).

  config 1: FatTree k=4, 64 flows between random host pairs, LogNormal(median 20 KB, sigma 1.5)
            sizes clipped to [1 KB, 10 MB], Poisson starts at 30% load; DCTCP, RED K=30 / 200
            packets, per-flow ECMP.
  config 2: FatTree k=16 as a topology file, 20,000 flows between random host pairs, sizes
            floor(Pareto(alpha 1.2, 1 KB)) capped at 30 MB, per-host Poisson starts at 50% load,
            25% of the flows in persistent connections of 2-8 flows (strictly increasing starts);
            per-packet ECMP (RngRun=1),
            FIN, the core s256 - aggregation s128 link down at 300 us and up at 700 us; NewReno,
            DCTCP and Cubic as separate worlds.
"""
import os

import numpy as np

LINK_BPS = 100e9


def fattree_topology(k):
    """The FatTree topology file text (same link order as the reference's topo.py)."""
    hk = k // 2
    H, E = k ** 3 // 4, k * hk
    links = [(f"h{h}", f"s{h // hk}") for h in range(H)]
    for pod in range(k):
        for a in range(hk):
            for e in range(hk):
                links.append((f"s{E + pod * hk + a}", f"s{pod * hk + e}"))
    for i in range(hk):
        for j in range(hk):
            for pod in range(k):
                links.append((f"s{2 * E + i * hk + j}", f"s{E + pod * hk + i}"))
    lines = [f"hosts {H}", f"switches {2 * E + hk * hk}"] + [f"link {a} {b}" for a, b in links]
    return "\n".join(lines) + "\n"


def _pairs(rng, n, H):
    src = rng.integers(0, H, n)
    dst = rng.integers(0, H - 1, n)
    dst = dst + (dst >= src)
    return src, dst


def config1_flows(seed=1, n=64, k=4, load=0.3):
    """64 lognormal flows; starts: a Poisson process at `load` of the hosts' capacity, first n."""
    rng = np.random.Generator(np.random.PCG64(seed))
    H = k ** 3 // 4
    size = np.clip(np.rint(rng.lognormal(np.log(20000.0), 1.5, n)), 1000, 10_000_000).astype(np.int64)
    mean = 20000.0 * np.exp(1.5 ** 2 / 2)  # the unclipped lognormal mean
    rate_per_ns = load * H * LINK_BPS / 8 / 1e9 / mean
    start = np.floor(np.cumsum(rng.exponential(1.0 / rate_per_ns, n))).astype(np.int64)
    src, dst = _pairs(rng, n, H)
    rows = sorted(zip(src.tolist(), dst.tolist(), size.tolist(), start.tolist()), key=lambda r: (r[3], r[0]))
    return "src,dst,flow_size,start_ns\n" + "".join(f"{a},{b},{c},{d}\n" for a, b, c, d in rows)


def config2_flows(seed=2, n=20000, k=16, load=0.5, alpha=1.2, xm=1000, cap=30_000_000, conn_share=0.25):
    """20,000 Pareto flows: per-host Poisson starts at `load`; a quarter of the flows are the later
    messages of persistent connections (2-8 messages each, one host pair, increasing starts)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    H = k ** 3 // 4
    # the mean of floor(Pareto) capped at `cap`, for the arrival rate
    mean = alpha * xm / (alpha - 1) * (1 - (xm / cap) ** (alpha - 1))
    rate_per_ns = load * LINK_BPS / 8 / 1e9 / mean
    per_host = n // H + 1
    gaps = rng.exponential(1.0 / rate_per_ns, (H, per_host))
    starts = np.floor(np.cumsum(gaps, axis=1)).astype(np.int64)
    # the n earliest (host, k-th arrival) events overall
    order = np.argsort(starts, axis=None, kind="stable")[:n]
    hosts, idx = np.unravel_index(order, starts.shape)
    src = hosts.astype(np.int64)
    start = starts[hosts, idx]
    size = np.minimum(np.floor(xm * (1 + rng.pareto(alpha, n))), cap).astype(np.int64)
    dst = rng.integers(0, H - 1, n)
    dst = dst + (dst >= src)
    conn = np.full(n, -1, dtype=np.int64)
    # persistent connections: runs of one host's consecutive flows share a connection (and its
    # destination) until a quarter of all flows are later messages of some connection
    by_host = {}
    for i in np.argsort(start, kind="stable"):
        by_host.setdefault(int(src[i]), []).append(int(i))
    later, cid = 0, 0
    host_list = sorted(by_host)
    rng.shuffle(host_list)
    for h in host_list:
        flows = by_host[h]
        pos = 0
        while later < conn_share * n and pos + 1 < len(flows):
            m = int(rng.integers(2, 9))
            group = flows[pos:pos + m]
            if len(group) < 2:
                break
            for j, i in enumerate(group):
                conn[i] = cid
                dst[i] = dst[group[0]]
                if j and start[i] <= start[group[j - 1]]:
                    # strictly later: a message written in the nanosecond its connection starts
                    # is dropped by the model (the start resets the due count, src/step.cuh:2391)
                    start[i] = start[group[j - 1]] + 1
            later += len(group) - 1
            cid += 1
            pos += m
        if later >= conn_share * n:
            break
    rows = sorted(zip(src.tolist(), dst.tolist(), size.tolist(), start.tolist(), conn.tolist()), key=lambda r: (r[3], r[0]))
    return "src,dst,flow_size,start_ns,conn\n" + "".join(f"{a},{b},{c},{d},{e}\n" for a, b, c, d, e in rows)


def leafspine_topology(leaves, spines, hpl):
    """Leaf-spine as a topology file (hosts on their leaf's first ports, then leaf-spine links)."""
    links = [(f"h{h}", f"s{h // hpl}") for h in range(leaves * hpl)]
    links += [(f"s{l}", f"s{leaves + j}") for l in range(leaves) for j in range(spines)]
    lines = [f"hosts {leaves * hpl}", f"switches {leaves + spines}"] + [f"link {a} {b}" for a, b in links]
    return "\n".join(lines) + "\n"


def jellyfish_topology(n, d, hps, seed):
    """A Jellyfish network (Singla et al., NSDI 2012): random d-regular switch graph built by
    joining random switch pairs with free ports that are not yet neighbours, breaking a random
    link to absorb a switch left with two or more free ports; hps hosts per switch."""
    rng = np.random.Generator(np.random.PCG64(seed))
    free = [d] * n
    adj = [set() for _ in range(n)]
    edges = []
    while True:
        avail = [x for x in range(n) if free[x] > 0]
        a = b = -1
        for _ in range(64 if len(avail) > 1 else 0):  # a random joinable pair, sampled while plentiful
            u, v = avail[int(rng.integers(len(avail)))], avail[int(rng.integers(len(avail)))]
            if u != v and v not in adj[u]:
                a, b = min(u, v), max(u, v)
                break
        if a < 0:
            pairs = [(u, v) for i, u in enumerate(avail) for v in avail[i + 1:] if v not in adj[u]]
            if pairs:
                a, b = pairs[int(rng.integers(len(pairs)))]
        if a < 0:
            stuck = [x for x in range(n) if free[x] >= 2]
            if not stuck:
                break
            x = stuck[int(rng.integers(len(stuck)))]
            cand = [e for e in edges if x not in e and e[0] not in adj[x] and e[1] not in adj[x]]
            if not cand:
                break
            y, z = cand[int(rng.integers(len(cand)))]
            edges.remove((y, z))
            adj[y].discard(z)
            adj[z].discard(y)
            free[y] += 1
            free[z] += 1
            for a, b in ((x, y), (x, z)):
                adj[a].add(b)
                adj[b].add(a)
                edges.append((min(a, b), max(a, b)))
                free[a] -= 1
                free[b] -= 1
            continue
        adj[a].add(b)
        adj[b].add(a)
        edges.append((a, b))
        free[a] -= 1
        free[b] -= 1
    H = n * hps
    links = [(f"h{h}", f"s{h // hps}") for h in range(H)] + [(f"s{a}", f"s{b}") for a, b in edges]
    return "\n".join([f"hosts {H}", f"switches {n}"] + [f"link {a} {b}" for a, b in links]) + "\n"


def torus_topology(x, y, hps):
    """A 2-D torus of x * y switches (4 network ports each), hps hosts per switch."""
    n = x * y
    sw = lambda i, j: (i % x) * y + (j % y)
    edges = []
    for i in range(x):
        for j in range(y):
            edges.append((sw(i, j), sw(i + 1, j)))
            edges.append((sw(i, j), sw(i, j + 1)))
    H = n * hps
    links = [(f"h{h}", f"s{h // hps}") for h in range(H)] + [(f"s{a}", f"s{b}") for a, b in edges]
    return "\n".join([f"hosts {H}", f"switches {n}"] + [f"link {a} {b}" for a, b in links]) + "\n"


def pareto_flows(seed, n, H, load=0.4, alpha=1.2, xm=1000, cap=30_000_000):
    """n Pareto-sized flows between random host pairs, per-host Poisson starts at `load`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mean = alpha * xm / (alpha - 1) * (1 - (xm / cap) ** (alpha - 1))
    rate_per_ns = load * LINK_BPS / 8 / 1e9 / mean
    per_host = n // H + 2
    starts = np.floor(np.cumsum(rng.exponential(1.0 / rate_per_ns, (H, per_host)), axis=1)).astype(np.int64)
    order = np.argsort(starts, axis=None, kind="stable")[:n]
    hosts, idx = np.unravel_index(order, starts.shape)
    src, start = hosts.astype(np.int64), starts[hosts, idx]
    size = np.minimum(np.floor(xm * (1 + rng.pareto(alpha, n))), cap).astype(np.int64)
    dst = rng.integers(0, H - 1, n)
    dst = dst + (dst >= src)
    rows = sorted(zip(src.tolist(), dst.tolist(), size.tolist(), start.tolist()), key=lambda r: (r[3], r[0]))
    return "src,dst,flow_size,start_ns\n" + "".join(f"{a},{b},{c},{d}\n" for a, b, c, d in rows)


def config4_worlds(outdir, n_worlds=256, seed=4):
    """The batched sweep: n_worlds worlds drawn by seed. Returns a list of `twins` argument lists
    (each with --flows, topology and parameters) and writes the files they name into outdir."""
    os.makedirs(outdir, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    worlds = []
    for w in range(n_worlds):
        kind = int(rng.integers(4))
        if kind == 0:  # FatTree
            k = int(rng.choice([4, 8, 16], p=[0.4, 0.4, 0.2]))
            topo, H = [f"--k={k}"], k ** 3 // 4
        elif kind == 1:  # leaf-spine
            leaves = int(rng.integers(8, 33))
            spines, hpl = int(rng.choice([4, 8])), 16
            tf = os.path.join(outdir, f"w{w:03d}_ls.txt")
            text = leafspine_topology(leaves, spines, hpl)
            topo, H = [f"--topo=leafspine", f"--leaves={leaves}", f"--spines={spines}", f"--hpl={hpl}"], leaves * hpl
        elif kind == 2:  # Jellyfish
            n = int(rng.integers(32, 257))
            d = int(rng.choice([4, 6, 8]))
            hps = int(rng.integers(2, 9))
            tf = os.path.join(outdir, f"w{w:03d}_jf.txt")
            text = jellyfish_topology(n, d, hps, seed * 1000 + w)
            topo, H = ["--topo=graph", f"--topofile={tf}"], n * hps
        else:  # torus
            x = int(rng.integers(6, 17))
            y = int(rng.integers(6, 17))
            hps = int(rng.integers(2, 9))
            tf = os.path.join(outdir, f"w{w:03d}_torus.txt")
            text = torus_topology(x, y, hps)
            topo, H = ["--topo=graph", f"--topofile={tf}"], x * y * hps
        if kind in (2, 3):
            if not os.path.exists(tf) or open(tf).read() != text:
                with open(tf, "w") as f:
                    f.write(text)
        queue = int(rng.choice([100, 200, 400]))
        ecn = int(rng.choice([k for k in (10, 30, 65, 100, 200) if k < queue]))
        tcp = str(rng.choice(["newreno", "dctcp", "cubic"]))
        nflows = int(rng.integers(1000, 20001))
        ff = os.path.join(outdir, f"w{w:03d}_flows.csv")
        text = pareto_flows(seed * 100000 + w, nflows, H)
        if not os.path.exists(ff) or open(ff).read() != text:
            with open(ff, "w") as f:
                f.write(text)
        worlds.append(topo + [f"--flows={ff}", f"--tcp={tcp}", f"--ecnK={ecn}", f"--queue={queue}"])
    return worlds


def config5_requests(outdir, n_requests=1000, seed=5, rate_per_s=100.0, ks=(4, 8, 16, 32), max_flows=64000, fail_share=0.25,
                     rdma_share=0.0, rdma_ks=(4, 8, 16)):
    """The serving loop's requests: Poisson arrivals at rate_per_s, each a FatTree world of k drawn
    from ks with U{1000..max_flows} Pareto-sized flows and a TCP variant drawn per request; a
    fail_share of them carry one link failure and recovery (then as a FatTree topology file).
    With rdma_share > 0 that share of the requests are RDMA worlds instead (`twins-rdma`: a FatTree
    topology file of k drawn from rdma_ks, the same kind of flow file, CC_MODE drawn from DCQCN, HPCC
    and DCTCP; no link failures: the RDMA model has none).
    Returns [(arrival_s, args)]: a `twins` argument list for TCP, a dict {topo, flows, overrides} for
    RDMA; writes the files into outdir."""
    os.makedirs(outdir, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    topo_files = {}

    def topo_file(k):
        if k not in topo_files:
            topo_files[k] = os.path.join(outdir, f"fattree{k}.txt")
            text = fattree_topology(k)
            if not os.path.exists(topo_files[k]) or open(topo_files[k]).read() != text:
                with open(topo_files[k], "w") as f:
                    f.write(text)
        return topo_files[k]

    for r in range(n_requests):
        t += float(rng.exponential(1.0 / rate_per_s))
        if rdma_share > 0 and rng.random() < rdma_share:
            k = int(rng.choice(rdma_ks))
            nflows = int(rng.integers(1000, max_flows + 1))
            cc = int(rng.choice([1, 3, 8]))
            ff = os.path.join(outdir, f"r{r:04d}_flows.csv")
            text = pareto_flows(seed * 1000003 + r, nflows, k ** 3 // 4)
            if not os.path.exists(ff) or os.path.getsize(ff) != len(text):
                with open(ff, "w") as f:
                    f.write(text)
            out.append((t, {"topo": topo_file(k), "flows": ff, "overrides": [f"CC_MODE={cc}"]}))
            continue
        k = int(rng.choice(ks))
        H = k ** 3 // 4
        nflows = int(rng.integers(1000, max_flows + 1))
        tcp = str(rng.choice(["newreno", "dctcp", "cubic"]))
        ff = os.path.join(outdir, f"r{r:04d}_flows.csv")
        text = pareto_flows(seed * 1000003 + r, nflows, H)
        if not os.path.exists(ff) or os.path.getsize(ff) != len(text):
            with open(ff, "w") as f:
                f.write(text)
        args = [f"--flows={ff}", f"--tcp={tcp}"]
        if rng.random() < fail_share:
            # a core-aggregation link of pod 0 down at 100-400 us, back up 300 us later
            core, agg = k * k // 2 + k * k // 2, k * k // 2
            down = int(rng.integers(100, 401)) * 1000
            args = ["--topo=graph", f"--topofile={topo_file(k)}", f"--fail={down}:s{core}:s{agg},{down + 300000}:s{core}:s{agg}:up"] + args
        else:
            args = [f"--k={k}"] + args
        out.append((t, args))
    return out


CONFIG2_ARGS = ["--topo=graph", "--spray=1", "--RngRun=1", "--fin=1", "--fail=300000:s256:s128,700000:s256:s128:up"]


def config3_flows(seed=3, leaves=32, hpl=16, load=0.4, duration_ns=1_000_000, median=50_000, sigma=1.5, incast_n=160,
                  incast_bytes=64 * 1024, incast_ns=500_000):
    """Config 3's flows on the 512-host leaf-spine: per-host Poisson all-to-all at `load` over
    `duration_ns` with LogNormal(median, sigma) sizes clipped to [1 KB, 10 MB], then a 160-to-1
    incast of 64 KB messages at 500 us from senders outside the receiver's leaf (the receiver drawn
    by the seed). Rows `src,dst,bytes,start_ns` after a header, background first (by start)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    H = leaves * hpl
    mean = median * np.exp(sigma ** 2 / 2)
    rate = load * LINK_BPS / 8 / mean / 1e9  # flows per ns per host
    rows = []
    for h in range(H):
        t = 0.0
        while True:
            t += rng.exponential(1 / rate)
            if t >= duration_ns:
                break
            d = int(rng.integers(0, H - 1))
            d += d >= h
            size = int(np.clip(round(rng.lognormal(np.log(median), sigma)), 1000, 10_000_000))
            rows.append((int(t), h, d, size))
    rows.sort()
    recv = int(rng.integers(0, H))
    others = [h for h in range(H) if h // hpl != recv // hpl]
    senders = rng.choice(others, size=incast_n, replace=False)
    lines = ["src,dst,flow_size,start_ns"] + [f"{a},{b},{n},{t}" for t, a, b, n in rows]
    lines += [f"{int(a)},{recv},{incast_bytes},{incast_ns}" for a in senders]
    return "\n".join(lines) + "\n"


def write_config3(outdir):
    """Writes config 3's topology and flows; returns (topology path, flows path, [(name, overrides)])."""
    os.makedirs(outdir, exist_ok=True)
    files = {"ls32.txt": leafspine_topology(32, 16, 16), "config3_flows.csv": config3_flows()}
    for name, text in files.items():
        p = os.path.join(outdir, name)
        if not os.path.exists(p) or open(p).read() != text:
            with open(p, "w") as f:
                f.write(text)
    return os.path.join(outdir, "ls32.txt"), os.path.join(outdir, "config3_flows.csv"), [("hpcc", ["CC_MODE=3"]), ("dcqcn", ["CC_MODE=1"])]


def write_configs(outdir):
    """Writes the config-1 and config-2 inputs into outdir; returns {name: (flows, args)}."""
    os.makedirs(outdir, exist_ok=True)
    p = lambda f: os.path.join(outdir, f)
    files = {"config1_flows.csv": config1_flows(), "config2_flows.csv": config2_flows(), "fattree16.txt": fattree_topology(16)}
    for name, text in files.items():
        if not os.path.exists(p(name)) or open(p(name)).read() != text:
            with open(p(name), "w") as f:
                f.write(text)
    return {
        "config1": (p("config1_flows.csv"), ["--k=4", "--tcp=dctcp", "--ecnK=30", "--queue=200"]),
        "config2": (p("config2_flows.csv"), CONFIG2_ARGS + [f"--topofile={p('fattree16.txt')}"]),
    }
