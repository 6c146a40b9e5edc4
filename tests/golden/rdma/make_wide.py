#!/usr/bin/env python3
"""Writes the wide-switch fixture: a leaf-spine world whose leaves have 44 ports (40 hosts and 4
spines each), more than the 32 that a switch's slot workspace holds, so that its leaves run the
serial path for switches of any width (rdma_model.h switch_step). The reference's CPU build cannot
run such a world (its switch workspaces hold 32 ports), so the expected flow completions are those
of this repository's CPU build of the model (tests/native/rdma_cpu), whose serial path equals the
reference on all of its validation jobs (tests/test_rdma_cpu.py runs them with
TWINS_RDMA_CPU_SERIAL=1). Flows (seed 7): a 40-to-1 incast of 256 KB at 20 us across the spines
plus 200 uniform flows of 1-200 KB over the first 100 us.

Runs after `make -C tests/native/rdma_cpu`: python tests/golden/rdma/make_wide.py"""
import os
import random
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
CPU = os.path.join(ROOT, "tests", "native", "rdma_cpu", "twins_rdma_cpu")
LEAVES, HPL, SPINES = 2, 40, 4
RUNS = [("cc1", ["CC_MODE=1"]), ("cc3", ["CC_MODE=3"]), ("cc8", ["CC_MODE=8"]), ("cc1_b3", ["CC_MODE=1", "BUFFER_SIZE=3"])]


def write_inputs():
    H = LEAVES * HPL
    topo = os.path.join(HERE, "inputs", "wide_ls2x40.topo")
    with open(topo, "w") as f:
        f.write(f"hosts {H}\nswitches {LEAVES + SPINES}\n")
        for h in range(H):
            f.write(f"link h{h} s{h // HPL}\n")
        for l in range(LEAVES):
            for p in range(SPINES):
                f.write(f"link s{l} s{LEAVES + p}\n")
    rng = random.Random(7)
    rows = []
    dst = 0
    for src in range(HPL, H):
        rows.append((src, dst, 262144, 20000))
    for _ in range(200):
        a, b = rng.sample(range(H), 2)
        rows.append((a, b, rng.randint(1000, 200000), rng.randint(0, 100000)))
    flows = os.path.join(HERE, "inputs", "wide_ls2x40.csv")
    with open(flows, "w") as f:
        f.write("src,dst,flow_size,start_ns\n")
        for r in rows:
            f.write("%d,%d,%d,%d\n" % r)
    return topo, flows


if __name__ == "__main__":
    topo, flows = write_inputs()
    for name, kv in RUNS:
        out = os.path.join(HERE, "ref", f"wide_ls2x40_{name}.fct")
        r = subprocess.run([CPU, os.path.join(HERE, "inputs", "hpcc.cfg"), f"--topo={topo}", f"--flows={flows}", f"--out={out}"] + kv,
                           capture_output=True, text=True, check=True)
        print(name, r.stdout.strip().splitlines()[-2:])
