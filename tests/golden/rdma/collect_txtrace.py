#!/usr/bin/env python3
"""Writes the transmission-trace fixtures: the reference's own RDMA model built for the CPU
(oracle/_ref/twins-rdma-cpu, see oracle/Makefile) run with --txtrace over a window of five of the
reference's validation jobs (PFC with DCQCN, HPCC with PFC, DCTCP-RDMA with 9 KB frames and a low
ECN threshold, leaf-spine, Jellyfish). Each fixture is the reference's trace file for that window:
"ns node dev kind src dst sport seq bytes ecn" per transmission start.

Runs where oracle/_ref is built: python tests/golden/rdma/collect_txtrace.py"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
BIN = os.path.join(ROOT, "oracle", "_ref", "twins-rdma-cpu")
JOBS = ["inc1_b1_cc1", "k8_incast16_b2_cc3", "inc1_b1_m9k_lowk_cc8", "ls_incast16_b2_cc1", "jf32_a2a30_cc3"]
WINDOW = (100000, 112000)  # ns, both ends included


def manifest():
    out = {}
    for line in open(os.path.join(HERE, "manifest.txt")):
        if line.startswith("#") or not line.strip():
            continue
        name, topo, flows, kv = line.split("|")[:4]
        out[name] = (topo, flows, kv.split())
    return out


if __name__ == "__main__":
    m = manifest()
    for name in JOBS:
        topo, flows, kv = m[name]
        out = os.path.join(HERE, "ref", f"tx_{name}.txt")
        subprocess.run([BIN, "hpcc.cfg", f"--topo={topo}", f"--flows={flows}", f"--out={out}.fct", f"--txtrace={out}",
                        f"--from={WINDOW[0]}", f"--until={WINDOW[1]}"] + kv, cwd=os.path.join(HERE, "inputs"), check=True,
                       capture_output=True)
        os.remove(out + ".fct")
        print(name, sum(1 for _ in open(out)), "records")
