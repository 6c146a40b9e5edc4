#!/usr/bin/env python3
"""Benchmark of the MI355X TCP packet-simulation core: the config-2 workload (the headline line)
and the config-4 batched sweep (the "sweep" object of the same JSON line).

A step is one complete simulation of the workload: FatTree k=16 (1024 hosts, 320 switches, as a
topology file), 20,000 Pareto-sized flows (a quarter in persistent connections), per-packet
ECMP, FIN, one link failure and recovery, run three times as separate worlds (NewReno, DCTCP,
Cubic) packed into one device batch. The inputs are uploaded before the timed region; the timed
region is the event loop (tw_batch_run), as the reference times its event loop.

Metric: simulated packet-events/s (a packet-event is one frame crossing one link), at bit-exact
flow times: every flow's establishment and completion time is compared with the oracle's in the
same run. roofline: the slot kernel's algorithmic HBM bytes (256 B per packet-event: a 64-byte
frame written and read through a link message and through an egress queue) over its HIP-event
time, against MI355X's 8 TB/s. cpu_baseline: the oracle (the sequential CPU restatement of the
model, one core) on the same three worlds.

Sweep (config 4): 256 worlds drawn by seed 4 (FatTree k in {4,8,16}, leaf-spine 8-32 leaves,
Jellyfish and torus graphs; ECN K, queue depth and TCP drawn per world; 1k-20k Pareto flows each),
one persistent launch per step through a HIP graph, with the launch plan (team sizes, queues) made
by the autotuner from the warm-up runs' per-world device times. Every world's flows are compared
with the oracle's; the plain per-world launch rate (each world alone, one after another) is
measured on every 8th world; the CPU baseline is the oracle on every world, on all the host cores
given, worlds in parallel.

RDMA (config 3, the "rdma" object): `twins-rdma` on the 512-host leaf-spine (32 leaves x 16
spines, 16 hosts per leaf, 100 Gb/s, 1 us), the reference's HPCC configuration, HPCC (CC_MODE 3)
and DCQCN (CC_MODE 1) as two worlds of one batch; background all-to-all LogNormal(50 KB, 1.5) at
40% load for 1 ms and a 160-to-1 incast of 64 KB at 500 us; port queues of 16,384 frames (the
default 4,096 overflows under the incast, which the reference reports as inexact). A step is one
complete run of both worlds. Parity: every flow's completion time against the reference's own RDMA
model built for the CPU (oracle/_ref, src/rdma compiled with TWINS_CPU), which is also the CPU
baseline (the two worlds on two cores at once, its event loop's wall time). roofline: 480 B per
packet-event (a 176-byte link frame written and read, a 64-byte queue entry written and read).

Reference workloads (the "reference_workloads" object): the reference's own single-world speed
table (data/twins_speed.csv, an RX 9060 XT) on the workloads whose ns-3 outputs are stored here,
each world alone through the plain launch, the event loop timed, every flow checked against ns-3;
and its twins-rdma table (README.md:324-333, the same GPU) on the validation jobs each row runs,
the event loop and the whole command timed, every flow checked against the reference's output.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--sweep 0|1] [--serve 0|1] [--rdma 0|1] [--refspeed 0|1]
"""
import argparse
import concurrent.futures as cf
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Hardware queues for this process's HIP streams (read when HIP starts, before any GPU call).
# With HIP's default of 4, the serving leg's six streams share queues, and work behind the TCP
# serving loop's persistent kernel on a shared queue waits until that kernel leaves: measured on
# the 1,000-request leg, an RDMA poll blocked for 15.8 s and the leg took 63 s (p99 53 s); with 16
# queues no host call took more than 0.4 s and the leg took 32 s (p99 21 s).
os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVENT = 256
TCPS = ("newreno", "dctcp", "cubic")


def oracle_run(flows, args, tcp, out):
    """The oracle on one world: (event-loop seconds, packet-events, fct rows)."""
    exe = os.path.join(ROOT, "oracle", "tcp_oracle")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    r = subprocess.run([exe] + args + [f"--flows={flows}", f"--tcp={tcp}", f"--out={out}"], capture_output=True, text=True, check=True)
    sec = float(re.search(r"Simulation Duration: ([0-9.]+)s", r.stdout).group(1))
    ev = int(re.search(r"packet-events (\d+)", r.stdout).group(1))
    rows = [tuple(int(x) for x in line.split(",")[4:6]) for line in open(out).read().splitlines()[1:]]
    return sec, ev, rows


def rdma_traffic_bytes():
    """HBM bytes of one config-3 step of the RDMA kernel, measured with rocprofv3 (see profiles/)."""
    p = os.path.join(ROOT, "profiles", "rdma_team_kernel_config3_traffic.json")
    return json.load(open(p))["traffic_bytes_per_step"] if os.path.exists(p) else None


def traffic_bytes():
    """The measured HBM bytes per launch (profiles/tcp_worlds_kernel_config2_traffic.json), or None."""
    p = os.path.join(ROOT, "profiles", "tcp_worlds_kernel_config2_traffic.json")
    try:
        return json.load(open(p))["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def oracle_world(args, out):
    """The oracle on one sweep world: (event-loop seconds, packet-events, fct rows)."""
    r = subprocess.run([os.path.join(ROOT, "oracle", "tcp_oracle")] + args + [f"--out={out}"], capture_output=True, text=True, check=True)
    sec = float(re.search(r"Simulation Duration: ([0-9.]+)s", r.stdout).group(1))
    ev = int(re.search(r"packet-events (\d+)", r.stdout).group(1))
    rows = [tuple(int(x) for x in line.split(",")[4:6]) for line in open(out).read().splitlines()[1:]]
    os.remove(out)
    return sec, ev, rows


def sweep_leg(a):
    """Config 4: the batched sweep; returns the "sweep" object of the JSON line."""
    import twins_amd
    from twins_amd import autotune, workloads
    wdir = os.path.join(a.sweep_dir, "worlds")
    t0 = time.time()
    wargs = workloads.config4_worlds(wdir, a.sweep_worlds, seed=4)
    worlds = [twins_amd.World(x) for x in wargs]
    log(f"sweep: {len(worlds)} worlds generated and built in {time.time() - t0:.1f} s")
    # plans of one workgroup per CU (the kernel's 256-thread workgroups fill a CU's registers)
    tuner = autotune.Autotuner([w.info for w in worlds], twins_amd.launch_capacity(9, 256), block_threads=256)
    for i in range(a.sweep_tune):
        plan = tuner.plan()
        t1 = time.time()
        b = twins_amd.Batch(worlds, plan=plan)
        t2 = time.time()
        st = b.run()
        tuner.observe(b)
        tuner.record(plan, st.wall_s)
        log(f"sweep tuning run {i}: teams {autotune.describe(plan)} pack {t2 - t1:.1f} s, run {st.wall_s:.3f} s, "
            f"{st.packet_events / st.wall_s:.3e} events/s, graph {st.graph}")
        del b
    plan = tuner.best()
    wall = kern = 0.0
    events = 0
    for i in range(a.steps):
        b = twins_amd.Batch(worlds, plan=plan)
        st = b.run()
        wall += st.wall_s
        kern += st.kernel_s
        events += st.packet_events
        log(f"sweep step {i}: {st.wall_s:.3f} s, {st.packet_events} packet-events, worlds done {st.worlds_done}")
        if i + 1 < a.steps:
            del b
    ws = [b.world_stats(w) for w in range(len(worlds))]
    top = sorted(range(len(worlds)), key=lambda w: -ws[w].device_s)[:5]
    for w in top:
        x = worlds[w].info
        log(f"  longest: world {w} ({x.hosts} hosts, {x.switches} switches, {x.flows} flows) team {ws[w].team_blocks}, "
            f"{ws[w].slots} slots, {ws[w].device_s:.3f} s ({1e6 * ws[w].device_s / max(1, ws[w].slots):.1f} us/slot)")
    # plain per-world launches (each world alone, the reference's one-world-per-launch form) on a sample
    sample = list(range(0, len(worlds), 8))
    p_wall = p_ev = 0.0
    for w in sample:
        st1 = twins_amd.Batch([worlds[w]]).run()
        p_wall += st1.wall_s
        p_ev += st1.packet_events
    log(f"sweep plain launches: {len(sample)} worlds, {p_wall:.2f} s")
    # the oracle on every world, all host cores: parity and the CPU baseline
    cores = a.cpu_cores
    t3 = time.time()
    with cf.ThreadPoolExecutor(cores) as ex:
        res = list(ex.map(lambda i: oracle_world(wargs[i], os.path.join(a.sweep_dir, f"o{i:03d}.csv")), range(len(worlds))))
    cpu_wall = time.time() - t3
    same = total = worlds_same = 0
    cpu_ev = 0
    for i, (sec, ev, rows) in enumerate(res):
        est, done = b.fct(i)
        s1 = sum(r == g for r, g in zip(rows, zip(est, done)))
        same += s1
        total += len(rows)
        worlds_same += s1 == len(rows) and b.counters(i)[26] == ev
        cpu_ev += ev
    log(f"sweep oracle: {cpu_wall:.1f} s on {cores} cores; parity {same}/{total}")
    value = events / wall
    achieved = BYTES_PER_EVENT * events / kern / 1e9
    return {
        "workload": f"config 4: {len(worlds)} worlds (seed 4) of FatTree k 4/8/16, leaf-spine 8-32 leaves, Jellyfish 32-256 and "
                    f"torus 36-256 switches; ECN K, queue, TCP drawn per world; 1k-20k Pareto(1.2) flows each; one persistent "
                    f"launch per step",
        "metric": "simulated packet-events/s at bit-exact FCT",
        "value": round(value, 1),
        "unit": "packet-events/s",
        "ms_per_step": round(1000 * wall / a.steps, 3),
        "steps": a.steps,
        "tuning_runs": a.sweep_tune,
        "packet_events_per_step": events // a.steps,
        "plan": {"teams": {str(k): v for k, v in autotune.describe(plan).items()}, "queues": len(set(plan["team_queue"]))},
        "graph": bool(st.graph),
        "parity": {"flows_identical_to_oracle": same, "flows": total, "worlds_identical": worlds_same, "worlds": len(worlds)},
        "plain_per_world": {"value": round(p_ev / p_wall, 1), "unit": "packet-events/s", "sample": f"every 8th world ({len(sample)}), "
                            "each alone in its own batch with the automatic launch shape, one after another"},
        "speedup_vs_plain": round(value / (p_ev / p_wall), 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "kernel": "tcp_worlds_kernel",
                     "kernel_ms_per_step": round(1000 * kern / a.steps, 3), "bytes_per_event": BYTES_PER_EVENT},
        "cpu_baseline": {"value": round(cpu_ev / cpu_wall, 1), "unit": "packet-events/s", "cores": cores, "kind": "port",
                         "sample": f"the oracle on all {len(worlds)} worlds, one process per world, {cores} at a time (wall time "
                                   "including process start and input parsing)"},
    }


def serve_leg(a):
    """Config 5: the serving loops, TCP and RDMA side by side; returns the "serve" object of the JSON line."""
    import twins_amd
    from twins_amd import rdma, workloads
    t0 = time.time()
    reqs = workloads.config5_requests(os.path.join(a.sweep_dir, "requests"), a.serve_requests, seed=5, rate_per_s=a.serve_rate,
                                      rdma_share=a.serve_rdma)
    cfg = os.path.join(ROOT, "tests", "golden", "rdma", "inputs", "hpcc.cfg")
    is_rdma = [isinstance(r[1], dict) for r in reqs]

    def build(i):
        r = reqs[i][1]
        return rdma.World(cfg, r["topo"], r["flows"], r["overrides"]) if is_rdma[i] else twins_amd.World(r)
    with cf.ThreadPoolExecutor(a.cpu_cores) as ex:
        worlds = list(ex.map(build, range(len(reqs))))
    n, n_rdma = len(reqs), sum(is_rdma)
    log(f"serve: {n} requests ({n_rdma} RDMA) generated and built in {time.time() - t0:.1f} s")
    # the GPU split: the RDMA loop's launches take rdma_blocks workgroups, the TCP loop's teams the rest
    resident = twins_amd.launch_capacity(9, 256)
    rblocks = a.serve_rdma_blocks if n_rdma else 0
    # (the TCP loop also leaves an eighth of the CUs free for the uploads' fill kernels)
    cus = resident  # (one slot-kernel workgroup per CU)
    nm, nl, zm, zl = a.serve_mid_teams, a.serve_large_teams, a.serve_mid_size, a.serve_large_size
    srv = twins_amd.Server(idle_s=2.0, mid_teams=nm, large_teams=nl, mid_size=zm, large_size=zl,
                           small_teams=resident - cus // 8 - rblocks - zl * nl - zm * nm)
    rsrv = rdma.Server(max_blocks=rblocks) if n_rdma else None
    # every serve_check-th request against the oracle (TCP) or the reference's RDMA build, on the
    # host's cores, each as soon as its result is in (the checks overlap the serving)
    ref = os.path.join(ROOT, "oracle", "_ref", "twins-rdma-cpu")

    def check(i, got):
        if is_rdma[i]:
            if not os.path.exists(ref):
                return i, None, got
            r = reqs[i][1]
            out = os.path.join(a.sweep_dir, f"s{i:04d}.fct")
            subprocess.run([ref, cfg, f"--topo={r['topo']}", f"--flows={r['flows']}", f"--out={out}", "--gwin=0"] + r["overrides"],
                           capture_output=True, timeout=1800)
            rows = open(out).read().splitlines()
            os.remove(out)
            return i, rows, got.splitlines()
        out = os.path.join(a.sweep_dir, f"s{i:04d}.csv")
        _, ev, rows = oracle_world(reqs[i][1], out)
        return i, rows, got
    pool = cf.ThreadPoolExecutor(max(1, a.cpu_cores - 2))
    checks = []
    sid = {}
    arrival, finish = {}, {}
    events = 0
    submitted = 0
    t_start = time.perf_counter()
    last_log = t_start
    inexact = 0
    slowest = [0.0, ""]  # the longest host call since the last progress line, and which

    def timed(what, fn, *args):
        c0 = time.perf_counter()
        r = fn(*args)
        d = time.perf_counter() - c0
        if d > slowest[0]:
            slowest[0], slowest[1] = d, what
        return r
    while len(finish) < n:
        now = time.perf_counter() - t_start
        while submitted < n and reqs[submitted][0] <= now:
            srv_ = rsrv if is_rdma[submitted] else srv
            sid[submitted] = timed("rdma submit" if is_rdma[submitted] else "tcp submit", srv_.submit, worlds[submitted])
            arrival[submitted] = reqs[submitted][0]
            submitted += 1
        st = timed("tcp poll", srv.poll)
        rst = timed("rdma poll", rsrv.poll) if rsrv else None
        for i in list(arrival):
            if i in finish:
                continue
            if is_rdma[i]:
                s_ = rsrv.state(sid[i])
                if s_ in (twins_amd.REQ_DONE, twins_amd.REQ_FAILED):
                    if s_ == twins_amd.REQ_FAILED:
                        raise RuntimeError(f"RDMA request {i} failed")
                    start, fct, ev = timed("rdma result", rsrv.result, sid[i])
                    finish[i] = time.perf_counter() - t_start
                    events += ev[38]
                    inexact += any(ev[j] for j in rdma.INEXACT)
                    if i % a.serve_check == 0:
                        checks.append(pool.submit(check, i, "".join(f"{j} {start[j]} {fct[j]}\n" for j in range(len(fct)) if fct[j] >= 0)))
            else:
                s_ = srv.state(sid[i])
                if s_ in (twins_amd.REQ_DONE, twins_amd.REQ_FAILED):
                    est, done, ev = timed("tcp result", srv.result, sid[i])
                    finish[i] = time.perf_counter() - t_start
                    events += ev[26]
                    if i % a.serve_check == 0:
                        checks.append(pool.submit(check, i, list(zip(est, done))))
        if time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            # (the TCP requests admitted and not yet taken, by their worlds' device footprint)
            held = sum(worlds[i].info.device_bytes for i in arrival if not is_rdma[i] and i not in finish and
                       srv.state(sid[i]) != twins_amd.REQ_WAITING)
            log(f"serve: {now:.0f} s, {submitted} submitted, {len(finish)} done; TCP resident {st.resident}, waiting {st.waiting}, "
                f"done {st.done}, relaunches {st.relaunches}, arena {st.arena_used / 2**30:.1f}/{st.arena_bytes / 2**30:.1f} GiB "
                f"(admitted worlds {held / 2**30:.1f} GiB)" +
                (f"; RDMA resident {rst.resident}, waiting {rst.waiting}, launches {rst.launches}, arena {rst.arena_used / 2**30:.1f}/"
                 f"{rst.arena_bytes / 2**30:.1f} GiB" if rst else "") + f"; slowest host call {slowest[1]} {slowest[0]:.3f} s")
            slowest[0], slowest[1] = 0.0, ""
        time.sleep(0.0005)
    wall = max(finish.values()) - reqs[0][0]
    srv.close()
    if rsrv:
        rsrv.close()
    lat = sorted(finish[r] - arrival[r] for r in finish)
    pct = lambda xs, p: xs[min(len(xs) - 1, int(p * len(xs)))]
    log(f"serve: {n} requests in {wall:.1f} s, p50 {pct(lat, 0.5):.3f} s, p99 {pct(lat, 0.99):.3f} s")
    by = {}
    for r in finish:
        by.setdefault(("rdma " if is_rdma[r] else "tcp ") + str(worlds[r].hosts if is_rdma[r] else worlds[r].info.hosts), []).append(
            finish[r] - arrival[r])
    for h, xs in sorted(by.items()):
        xs.sort()
        log(f"  {h} hosts: {len(xs)} requests, p50 {xs[len(xs) // 2]:.3f} s, max {xs[-1]:.3f} s")
    kind_lat = {}
    for kind in (False, True):
        xs = sorted(finish[r] - arrival[r] for r in finish if is_rdma[r] == kind)
        if xs:
            kind_lat["rdma" if kind else "tcp"] = {"requests": len(xs), "p50_s": round(pct(xs, 0.5), 4), "p99_s": round(pct(xs, 0.99), 4)}
    t_chk = time.time()
    same = total = checked = 0
    for f in checks:
        i, rows, got = f.result()
        if rows is None:
            continue
        checked += 1
        same += sum(x == y for x, y in zip(rows, got)) if len(rows) == len(got) else 0
        total += len(rows)
    pool.shutdown()
    log(f"serve: parity {same}/{total} flows of {checked} requests checked ({time.time() - t_chk:.1f} s after the last result)")
    return {
        "workload": f"config 5: {n} requests arriving Poisson at {a.serve_rate:g}/s (seed 5); TCP requests: a FatTree k in "
                    f"{{4,8,16,32}} world with U{{1000..64000}} Pareto(1.2) flows and a TCP variant drawn per request, a quarter with "
                    f"a link failure and recovery; RDMA requests ({n_rdma}, a share of {a.serve_rdma:g}): FatTree k in {{4,8,16}} "
                    f"with the same kind of flows, CC_MODE drawn from DCQCN, HPCC, DCTCP; two serving loops side by side on the "
                    f"GPU (the RDMA loop's launches on {rblocks} workgroups, the TCP loop's teams on the rest)",
        "metric": "sweep-request p50/p99 latency (arrival to flow times ready) and packet-events/s",
        "p50_s": round(pct(lat, 0.5), 4),
        "p99_s": round(pct(lat, 0.99), 4),
        "by_kind": kind_lat,
        "value": round(events / wall, 1),
        "unit": "packet-events/s",
        "requests": n,
        "wall_s": round(wall, 3),
        "rdma_inexact_requests": inexact,
        "parity": {"flows_identical_to_oracle": same, "flows": total, "requests_checked": checked,
                   "oracle": "TCP: oracle/tcp_oracle; RDMA: the reference's RDMA build (oracle/_ref)"},
    }


# the reference's own speed table (data/twins_speed.csv: its event loop's wall time on an RX 9060 XT,
# min of 3 runs, one world per process) for the workloads whose ns-3 outputs are stored here
REF_SPEED = [("k8_paper", "dctcp", 8, 0.2238), ("k8_paper", "newreno", 8, 0.3648), ("k8_a2a30", "dctcp", 8, 0.3262),
             ("k8_a2a30", "newreno", 8, 0.396), ("k8_a2a50s", "dctcp", 8, 0.357), ("k16_paper", "dctcp", 16, 0.434),
             ("k16_a2a30", "dctcp", 16, 0.6371), ("k16_a2a50s", "dctcp", 16, 0.5073), ("k16_a2a50s", "newreno", 16, 1.0502)]
TCP_FILE = {"dctcp": "TcpDctcp", "newreno": "TcpNewReno", "cubic": "TcpCubic"}


def refspeed_leg(a):
    """The reference's published single-world workloads: each world alone (the drop-in's plain
    launch, automatic shape), its event loop timed (tw_batch_run, min of a.ref_runs), every flow
    checked against the stored ns-3 output; returns the "reference_workloads" object."""
    import twins_amd
    g = os.path.join(ROOT, "tests", "golden")
    out = {}
    for name, tcp, k, ref_s in REF_SPEED:
        w = twins_amd.World([f"--k={k}", f"--flows={g}/flows/{name}.csv", f"--tcp={tcp}"])
        best, b, runs = None, None, 9
        for _ in range(a.ref_runs):
            if b is not None:
                b.close()
            b = twins_amd.Batch([w], chain_runs=runs)
            st = b.run()
            if b.counters(0)[25] and runs == 9:
                runs, best = 25, None  # undecided ties: 25-run chains from here on (the reference's twins128)
                b.close()
                b = twins_amd.Batch([w], chain_runs=runs)
                st = b.run()
            best = st.wall_s if best is None else min(best, st.wall_s)
        est, done = b.fct(0)
        exp = []
        with open(f"{g}/ns3/{name}_{TCP_FILE[tcp]}.csv") as f:
            next(f)
            exp = [tuple(int(x) for x in line.split(",")[4:6]) for line in f]
        same = sum(e == x for e, x in zip(exp, zip(est, done)))
        slots = b.counters(0)[31]
        b.close()
        out[f"{name}/{tcp}"] = {"event_loop_s": round(best, 4), "reference_rx9060xt_s": ref_s, "speedup_vs_reference": round(ref_s / best, 3),
                                "slots": slots, "us_per_slot": round(1e6 * best / max(1, slots), 2), "chain_runs": runs,
                                "flows_identical_to_ns3": same, "flows": len(exp)}
        log(f"refspeed: {name} {tcp}: {best:.4f} s ({slots} slots, {1e6 * best / max(1, slots):.1f} us/slot), reference "
            f"{ref_s} s on an RX 9060 XT; {same}/{len(exp)} flows identical to ns-3")
    out.update(refspeed_rdma(a))
    return out


# the reference's twins-rdma speed table (README.md:324-333: wall times on an RX 9060 XT), by the
# validation job each row runs
REF_SPEED_RDMA = [("k16_a2a30_cc1", 0.99), ("k16_a2a30_cc3", 1.14), ("k16_a2a50s_b2_cc1", 1.17), ("k8_a2a50s_cc3", 0.44),
                  ("k8_perm_cc3", 0.094), ("ls_a2a30_cc1", 0.63), ("jf32_a2a30_cc3", 0.76), ("k8_incast64_b1_cc1", 0.92)]


def refspeed_rdma(a):
    """The reference's published twins-rdma runs: each job alone, the event loop timed (trd_batch_run,
    min of a.ref_runs) and the whole `twins-rdma` command (set-up, upload, event loop, output) timed
    once, as the reference's table times its runs; every flow checked against the reference's output."""
    from twins_amd import rdma
    rg = os.path.join(ROOT, "tests", "golden", "rdma")
    inp, cfg = os.path.join(rg, "inputs"), os.path.join(rg, "inputs", "hpcc.cfg")
    jobs = {}
    for line in open(os.path.join(rg, "manifest.txt")):
        if not line.startswith("#") and line.strip():
            name, topo, flows, kv, _ = line.split("|", 4)
            jobs[name] = (topo, flows, kv.split())
    out = {}
    for name, ref_s in REF_SPEED_RDMA:
        topo, flows, kv = jobs[name]
        w = rdma.World(cfg, f"{inp}/{topo}", f"{inp}/{flows}", kv)
        best = None
        for _ in range(a.ref_runs):
            b = rdma.Batch([w])
            st = b.run()
            best = st.wall_s if best is None else min(best, st.wall_s)
            got = b.fct_text(0)
            b.close()
        same = got == open(os.path.join(rg, "ref", name + ".fct")).read()
        o = f"/tmp/twins_refspeed_{name}.fct"
        t0 = time.perf_counter()
        r = subprocess.run([os.path.join(ROOT, "twins_amd", "bin", "twins-rdma"), cfg, f"--topo={inp}/{topo}", f"--flows={inp}/{flows}",
                            f"--out={o}"] + kv, capture_output=True, text=True)
        cli = time.perf_counter() - t0
        same = same and r.returncode == 0 and open(o).read() == got
        out[f"rdma:{name}"] = {"event_loop_s": round(best, 4), "command_wall_s": round(cli, 4), "reference_rx9060xt_s": ref_s,
                               "speedup_vs_reference": round(ref_s / cli, 3), "flows_identical_to_reference": same}
        log(f"refspeed: rdma {name}: event loop {best:.4f} s, twins-rdma command {cli:.3f} s, reference {ref_s} s on an RX 9060 XT; "
            f"{'identical to' if same else 'DIFFERENT from'} the reference's output")
    return out


RDMA_BYTES_PER_EVENT = 480
RDMA_Q3CAP = 16384


def rdma_leg(a):
    """Config 3: the RDMA model on the 512-host leaf-spine; returns the "rdma" object of the JSON line."""
    from twins_amd import rdma, workloads
    topo, flows, runs = workloads.write_config3(os.path.join(ROOT, "tests", "golden", "configs"))
    cfg = os.path.join(ROOT, "tests", "golden", "rdma", "inputs", "hpcc.cfg")

    def step():
        ws = [rdma.World(cfg, topo, flows, kv) for _, kv in runs]
        b = rdma.Batch(ws, q3cap=RDMA_Q3CAP)
        st = b.run()
        return ws, b, st

    log(f"rdma: config 3, {len(runs)} worlds, {a.warmup} warm-up + {a.steps} steps")
    for _ in range(a.warmup):
        step()
    wall, kern, events = 0.0, 0.0, 0
    for _ in range(a.steps):
        ws, b, st = step()
        wall += st.wall_s
        kern += st.kernel_s
        events += b.counters()[38]
    ev = b.counters()
    inexact = [rdma.EV_NAMES[i] for i in rdma.INEXACT if ev[i]]
    per_step = events // a.steps
    # the reference's RDMA model built for the CPU: parity target and CPU baseline (both worlds at once)
    ref = os.path.join(ROOT, "oracle", "_ref", "twins-rdma-cpu")
    same, total, cpu = 0, 0, None
    if os.path.exists(ref):
        procs = []
        for name, kv in runs:
            out = f"/tmp/bench_rdma_ref_{name}.fct"
            procs.append((out, subprocess.Popen([ref, cfg, f"--topo={topo}", f"--flows={flows}", f"--out={out}", "--gwin=0",
                                                 f"--q3cap={RDMA_Q3CAP}"] + kv, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        walls = []
        for i, (out, p) in enumerate(procs):
            so, se = p.communicate(timeout=600)
            m = re.search(r"in ([0-9.]+) s wall", so)
            walls.append(float(m.group(1)) if m else float("nan"))
            want = open(out).read().splitlines()
            got = b.fct_text(i).splitlines()
            same += sum(x == y for x, y in zip(want, got)) if len(want) == len(got) else 0
            total += ws[i].flows
            log(f"rdma: reference CPU build, {runs[i][0]}: {walls[-1]:.2f} s; {len(want)} flows done")
        cpu = {"value": round(per_step / max(walls), 1), "unit": "packet-events/s", "cores": len(runs), "kind": "reference",
               "sample": "the reference's RDMA model built for the CPU (oracle/_ref), the two worlds of one step as two processes "
                         "at once, its event loop's wall time (the longer of the two)"}
    else:
        log("rdma: oracle/_ref/twins-rdma-cpu is not built: no parity check or CPU baseline")
    achieved = RDMA_BYTES_PER_EVENT * events / kern / 1e9
    return {
        "metric": "simulated packet-events/s at bit-exact FCT",
        "value": round(events / wall, 1),
        "unit": "packet-events/s",
        "steps": a.steps,
        "ms_per_step": round(1000 * wall / a.steps, 3),
        "config": {"workload": "config 3: twins-rdma, leaf-spine 32 x 16 (512 hosts, 48 switches of 32 ports), 100 Gb/s 1 us; "
                               "16,768 LogNormal(50 KB, 1.5) all-to-all flows at 40% load over 1 ms + 160-to-1 incast of 64 KB at "
                               "500 us; HPCC (CC_MODE 3) and DCQCN (CC_MODE 1) as 2 worlds of one batch; q3cap 16384",
                   "worlds": len(runs), "packet_events_per_step": per_step, "slots_max": st.slots, "flows_done": st.flows_done,
                   "inexact_counters": inexact},
        "parity": {"flows_identical_to_reference": same, "flows": total,
                   "reference": "src/rdma (the reference's model) compiled for the CPU: oracle/_ref/twins-rdma-cpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": rdma_traffic_bytes(),
                     "traffic_source": "HBM bytes per step (2 launches: a launch ends after 1 s) of this kernel on this workload: "
                                       "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy's readings "
                                       "(profiles/rdma_team_kernel_config3.md)",
                     "kernel": "rdma_team_kernel",
                     "kernel_ms_per_step": round(1000 * kern / a.steps, 3), "bytes_per_event": RDMA_BYTES_PER_EVENT},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--team_blocks", type=int, default=0)
    ap.add_argument("--hosts_per_wave", type=int, default=0)
    ap.add_argument("--block_threads", type=int, default=0)
    ap.add_argument("--sweep", type=int, default=1, help="also run the config-4 batched sweep (the \"sweep\" object)")
    ap.add_argument("--sweep_worlds", type=int, default=256)
    ap.add_argument("--sweep_tune", type=int, default=3, help="autotuning runs before the timed sweep steps")
    ap.add_argument("--sweep_dir", default="/tmp/twins_sweep")
    ap.add_argument("--cpu_cores", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--serve", type=int, default=1, help="also run the config-5 serving loop (the \"serve\" object)")
    ap.add_argument("--serve_requests", type=int, default=1000)
    ap.add_argument("--serve_rate", type=float, default=100.0)
    ap.add_argument("--serve_check", type=int, default=1, help="every n-th request is checked against the oracle")
    ap.add_argument("--serve_rdma", type=float, default=0.1, help="the share of the serving loop's requests that are RDMA worlds")
    ap.add_argument("--serve_rdma_blocks", type=int, default=64, help="workgroups the RDMA serving loop's launches take")
    ap.add_argument("--serve_mid_teams", type=int, default=8, help="the TCP serving loop's mid teams (FatTree k=16 worlds)")
    ap.add_argument("--serve_large_teams", type=int, default=4, help="the TCP serving loop's large teams (k=32 worlds)")
    ap.add_argument("--serve_mid_size", type=int, default=4, help="workgroups of the TCP serving loop's mid teams")
    ap.add_argument("--serve_large_size", type=int, default=16, help="workgroups of the TCP serving loop's large teams")
    ap.add_argument("--rdma", type=int, default=1, help="also run config 3, the RDMA model (the \"rdma\" object)")
    ap.add_argument("--refspeed", type=int, default=1, help="also time the reference's published single-world workloads")
    ap.add_argument("--ref_runs", type=int, default=3, help="runs per reference workload (the best is reported)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU (the worlds of a sweep are independent and would be "
              f"spread over GPUs by HBM footprint); --gpus {a.gpus} is not run", file=sys.stderr)
        sys.exit(2)
    import twins_amd
    from twins_amd import workloads
    cfg = workloads.write_configs(os.path.join(ROOT, "tests", "golden", "configs"))
    flows, args = cfg["config2"]
    worlds = [twins_amd.World(args + [f"--flows={flows}", f"--tcp={t}"]) for t in TCPS]

    def step():
        b = twins_amd.Batch(worlds, team_blocks=a.team_blocks, hosts_per_wave=a.hosts_per_wave, block_threads=a.block_threads)
        st = b.run()
        return b, st

    for _ in range(a.warmup):
        step()
    wall, kern, events, runs = 0.0, 0.0, 0, []
    for _ in range(a.steps):
        b, st = step()
        wall += st.wall_s
        kern += st.kernel_s
        events += st.packet_events
        runs.append((b, st))
    b, st = runs[-1]
    # bit-exact check against the oracle, and the oracle's timing (one core)
    cpu_sec, cpu_ev, same, total = 0.0, 0, 0, 0
    for i, t in enumerate(TCPS):
        sec, ev, rows = oracle_run(flows, args, t, f"/tmp/bench_oracle_{t}.csv")
        cpu_sec += sec
        cpu_ev += ev
        est, done = b.fct(i)
        same += sum(r == (e, d) for r, (e, d) in zip(rows, zip(est, done)))
        total += len(rows)
        assert b.counters(i)[26] == ev, f"{t}: packet-events differ from the oracle's"
    value = events / wall
    achieved = BYTES_PER_EVENT * events / kern / 1e9
    line = {
        "metric": "simulated packet-events/s at bit-exact FCT",
        "value": round(value, 1),
        "unit": "packet-events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000 * wall / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "config 2: FatTree k=16 topology file (1024 hosts, 320 switches), 20,000 Pareto(1.2, 1 KB) flows "
                        "(25% in persistent connections), spray (RngRun=1), FIN, link s256-s128 down 300-700 us; "
                        "NewReno + DCTCP + Cubic as 3 worlds of one batch",
            "worlds": 3,
            "packet_events_per_step": events // a.steps,
            "slots_max": st.slots_max,
            "flows_done": st.flows_done,
            "parallelism": "one device batch; teams of workgroups per world",
        },
        "parity": {"flows_identical_to_oracle": same, "flows": total},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic_bytes(),
            "traffic_source": "HBM bytes per launch of this kernel on this workload: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                              "runs, scaled by the same counters' readings for a 1 GiB copy (profiles/tcp_worlds_kernel_config2.md)",
            "kernel": "tcp_worlds_kernel",
            "kernel_ms_per_step": round(1000 * kern / a.steps, 3),
            "bytes_per_event": BYTES_PER_EVENT,
        },
        "cpu_baseline": {
            "value": round(cpu_ev / cpu_sec, 1),
            "unit": "packet-events/s",
            "cores": 1,
            "kind": "port",
            "sample": "the oracle (sequential CPU restatement of the model) on the same 3 worlds, event loop only",
        },
    }
    if a.sweep:
        line["sweep"] = sweep_leg(a)
    if a.serve:
        line["serve"] = serve_leg(a)
    if a.rdma:
        line["rdma"] = rdma_leg(a)
    if a.refspeed:
        line["reference_workloads"] = refspeed_leg(a)
    print(json.dumps(line), flush=True)
    bad = [k for k in ("sweep", "serve") if k in line and line[k]["parity"]["flows_identical_to_oracle"] != line[k]["parity"]["flows"]]
    if "rdma" in line:
        r = line["rdma"]
        if r["parity"]["flows_identical_to_reference"] != r["parity"]["flows"] or r["config"]["inexact_counters"]:
            bad.append("rdma")
    if "reference_workloads" in line:
        bad += [k for k, v in line["reference_workloads"].items()
                if (v["flows_identical_to_ns3"] != v["flows"] if "flows" in v else not v["flows_identical_to_reference"])]
    if same != total or bad:
        sys.exit(1)


if __name__ == "__main__":
    main()
