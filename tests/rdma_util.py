"""Helpers of the RDMA parity tests: the reference's RDMA jobs stored under tests/golden/rdma (the
reference's own RDMA model run on the CPU over its validation job files; see collect_rdma.py there)."""
import os

import util

RG = os.path.join(util.GOLDEN, "rdma")
INPUTS = os.path.join(RG, "inputs")
CONFIG = os.path.join(INPUTS, "hpcc.cfg")
CPU_BIN = os.path.join(util.ROOT, "tests", "native", "rdma_cpu", "twins_rdma_cpu")


def jobs(max_ref_s=None, min_ref_s=None):
    """(name, topology, flows, overrides, reference CPU seconds, reference counter line) per job."""
    out = []
    for line in open(os.path.join(RG, "manifest.txt")):
        if line.startswith("#") or not line.strip():
            continue
        name, topo, flows, kv, sec, summ = line.rstrip("\n").split("|", 5)
        sec = float(sec)
        if (max_ref_s is not None and sec > max_ref_s) or (min_ref_s is not None and sec <= min_ref_s):
            continue
        counters = summ.split(" ; ")[-1]
        out.append((name, topo, flows, kv.split(), sec, counters))
    return out


def ref_fct(name):
    return open(os.path.join(RG, "ref", name + ".fct")).read()


def sum_counters(lines):
    """The element-wise sum of counter lines (a batch's counters are its worlds' sums)."""
    import re
    tot = None
    for l in lines:
        v = [int(x) for x in re.findall(r"\d+", l)]
        tot = v if tot is None else [a + b for a, b in zip(tot, v)]
    return tot


# the transmission-trace fixtures (collect_txtrace.py): the reference's --txtrace files of these jobs
# over this window (ns, both ends included)
TX_WINDOW = (100000, 112000)


def tx_jobs():
    names = sorted(f[3:-4] for f in os.listdir(os.path.join(RG, "ref")) if f.startswith("tx_") and f.endswith(".txt"))
    by = {j[0]: j for j in jobs()}
    return [by[n] for n in names]


def ref_txtrace(name):
    return open(os.path.join(RG, "ref", f"tx_{name}.txt"), "rb").read()


def wide_runs():
    """The wide-switch fixture (make_wide.py): (name, overrides) per run of its world."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_wide", os.path.join(RG, "make_wide.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.RUNS


WIDE_TOPO, WIDE_FLOWS = "wide_ls2x40.topo", "wide_ls2x40.csv"
