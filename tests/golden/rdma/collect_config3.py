#!/usr/bin/env python3
"""Writes the config-3 RDMA fixtures: the reference's own RDMA model built for the CPU
(oracle/_ref/twins-rdma-cpu, see oracle/Makefile) run on the config-3 inputs
(tests/golden/configs/ls32.txt and config3_flows.csv, written by twins_amd.workloads.write_config3)
with the reference's HPCC configuration, HPCC and DCQCN, port queues of 16,384 frames. Keeps each
run's flow completions (`id start_ns fct_ns`) and its counter line.

Runs where oracle/_ref is built: python tests/golden/rdma/collect_config3.py"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, ROOT)

from twins_amd import workloads  # noqa: E402

BIN = os.path.join(ROOT, "oracle", "_ref", "twins-rdma-cpu")

if __name__ == "__main__":
    topo, flows, runs = workloads.write_config3(os.path.join(ROOT, "tests", "golden", "configs"))
    with open(os.path.join(HERE, "config3_manifest.txt"), "w") as man:
        man.write("# run|overrides|reference counter line (q3cap 16384)\n")
        for name, kv in runs:
            out = os.path.join(HERE, "ref", f"config3_{name}.fct")
            r = subprocess.run([BIN, os.path.join(HERE, "inputs", "hpcc.cfg"), f"--topo={topo}", f"--flows={flows}", f"--out={out}",
                                "--gwin=0", "--q3cap=16384"] + kv, capture_output=True, text=True, check=True)
            line = [l for l in r.stdout.splitlines() if l.startswith("drops ")][0]
            man.write(f"{name}|{' '.join(kv)}|{line}\n")
            print(name, line)
