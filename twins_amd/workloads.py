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


CONFIG2_ARGS = ["--topo=graph", "--spray=1", "--RngRun=1", "--fin=1", "--fail=300000:s256:s128,700000:s256:s128:up"]


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
