#!/usr/bin/env python3
"""Benchmark of the MI355X TCP packet-simulation core on the config-2 workload.

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

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--team_blocks", type=int, default=0)
    ap.add_argument("--hosts_per_wave", type=int, default=0)
    ap.add_argument("--block_threads", type=int, default=0)
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
            "traffic": None,
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
    print(json.dumps(line), flush=True)
    if same != total:
        sys.exit(1)


if __name__ == "__main__":
    main()
