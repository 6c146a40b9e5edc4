#!/usr/bin/env python3
"""Writes the RDMA golden fixtures: every job of the reference's RDMA validation job files
(validation/hpcc/jobs1.txt, jobs2.txt, jobs3.txt, jobs_dctcp.txt, sweep_k8.txt; each line
`name topo flows [KEY=VALUE ...]`) run through the reference's own RDMA model compiled for the CPU
(oracle/_ref/twins-rdma-cpu, see oracle/Makefile), with the reference's HPCC configuration
(validation/hpcc/hpcc.cfg). For each job it keeps the flow completions (`id start_ns fct_ns`,
the reference's fct output) and the summary counters, and copies the topology and flow files the
jobs name into inputs/. The reference's validation states that these jobs are identical to the
HPCC simulator (validation/hpcc/README.md:60-75).

Runs only where /root/reference exists: python tests/golden/rdma/collect_rdma.py [max_seconds]"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
REF = "/root/reference"
BIN = os.path.join(ROOT, "oracle", "_ref", "twins-rdma-cpu")
JOBS = ["jobs1.txt", "jobs2.txt", "jobs3.txt", "jobs_dctcp.txt", "sweep_k8.txt"]


def find(name):
    for d in ("data/rdma", "data/topo", "data/flows"):
        p = os.path.join(REF, d, name)
        if os.path.exists(p):
            return p
    raise FileNotFoundError(name)


def run(job):
    name, topo, flows, extra, src = job
    out = os.path.join(HERE, "ref", name + ".fct")
    t0 = time.time()
    r = subprocess.run([BIN, os.path.join(REF, "validation/hpcc/hpcc.cfg"), f"--topo={find(topo)}", f"--flows={find(flows)}",
                        f"--out={out}", "--gwin=0"] + extra, capture_output=True, text=True, timeout=LIMIT)
    summary = [l for l in r.stdout.splitlines() if l.startswith("drops ") or l.startswith("flows done")]
    return name, r.returncode, time.time() - t0, summary


if __name__ == "__main__":
    LIMIT = float(sys.argv[1]) if len(sys.argv) > 1 else 600
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(os.path.join(HERE, "ref"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    jobs = []
    for jf in JOBS:
        for line in open(os.path.join(REF, "validation/hpcc", jf)):
            f = line.split()
            if not f or f[0].startswith("#"):
                continue
            jobs.append((f[0], f[1], f[2], f[3:], jf))
    manifest = []
    with cf.ThreadPoolExecutor(8) as ex:
        futs = {ex.submit(run, j): j for j in jobs}
        for fu in cf.as_completed(futs):
            j = futs[fu]
            try:
                name, rc, sec, summary = fu.result()
            except subprocess.TimeoutExpired:
                print(f"{j[0]}: over {LIMIT:.0f} s, left out", flush=True)
                continue
            print(f"{name}: rc {rc} {sec:.1f} s {' | '.join(summary)}", flush=True)
            if rc != 0:
                continue
            for x in (j[1], j[2]):
                dst = os.path.join(HERE, "inputs", x)
                if not os.path.exists(dst):
                    shutil.copy(find(x), dst)
            manifest.append((j, sec, summary))
    manifest.sort(key=lambda m: (JOBS.index(m[0][4]), m[0][0]))
    with open(os.path.join(HERE, "manifest.txt"), "w") as f:
        f.write("# job|topology|flows|KEY=VALUE overrides|reference seconds (CPU)|reference summary\n")
        for (name, topo, flows, extra, src), sec, summary in manifest:
            f.write(f"{name}|{topo}|{flows}|{' '.join(extra)}|{sec:.1f}|{' ; '.join(summary)}\n")
    print(f"{len(manifest)} of {len(jobs)} jobs written")
