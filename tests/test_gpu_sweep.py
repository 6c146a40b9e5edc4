"""Config 4 on the GPU: the batched sweep's own workload (256 worlds drawn by seed 4: FatTree k 4/8/16,
leaf-spine, Jellyfish and torus graphs; every ECN K, queue depth and TCP variant of the draw) run
as one persistent launch through the HIP graph with an autotuned plan. Every world's flow times
and packet-event count must equal the oracle's; a world whose 9-run chains leave ties undecided is
run again in a batch of 25-run chains, as the reference reruns with twins128."""
import concurrent.futures as cf
import os
import re
import subprocess

import pytest

import util


def world_features(args):
    kind = "fattree" if any(a.startswith("--k=") for a in args) else "leafspine" if "--topo=leafspine" in args else \
        "jellyfish" if any(a.endswith("_jf.txt") for a in args) else "torus"
    kv = dict(a[2:].split("=", 1) for a in args if a.startswith("--") and "=" in a)
    return kind, int(kv["ecnK"]), int(kv["queue"]), kv["tcp"]


def oracle(args, out, runs=9):
    r = subprocess.run([util.oracle_bin()] + args + [f"--out={out}", f"--chain_runs={runs}"], capture_output=True, text=True,
                       check=True)
    ev = int(re.search(r"packet-events (\d+)", r.stdout).group(1))
    return util.read_cols(out, (4, 5)), ev


@pytest.mark.gpu
def test_gpu_config4_sweep(tmp_path):
    import twins_amd
    from twins_amd import autotune, workloads
    wargs = workloads.config4_worlds(str(tmp_path / "worlds"), 256, seed=4)
    feats = [world_features(a) for a in wargs]
    # the draw covers every kind, ECN K, queue depth and TCP variant
    assert {f[0] for f in feats} == {"fattree", "leafspine", "jellyfish", "torus"}
    assert {f[1] for f in feats} == {10, 30, 65, 100, 200} and {f[2] for f in feats} == {100, 200, 400}
    assert {f[3] for f in feats} == {"newreno", "dctcp", "cubic"}
    worlds = [twins_amd.World(a) for a in wargs]
    print(f"config 4: {len(worlds)} worlds", flush=True)
    tuner = autotune.Autotuner([w.info for w in worlds], twins_amd.launch_capacity(9, 256), block_threads=256)
    b = twins_amd.Batch(worlds, plan=tuner.plan())
    st = b.run()
    tuner.observe(b)
    b.close()
    plan = tuner.plan()
    print(f"tuned plan: teams {autotune.describe(plan)}", flush=True)
    b = twins_amd.Batch(worlds, plan=plan)
    st = b.run()
    assert st.graph and st.worlds_done == len(worlds)
    cores = min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(cores) as ex:
        ref = list(ex.map(lambda i: oracle(wargs[i], str(tmp_path / f"o{i:03d}.csv")), range(len(worlds))))
    tied, bad = [], []
    for i, (rows, ev) in enumerate(ref):
        est, done = b.fct(i)
        got = [(str(e), str(d)) for e, d in zip(est, done)]
        if b.counters(i)[25]:
            tied.append(i)
            continue
        if got != rows or b.counters(i)[26] != ev:
            bad.append(i)
    b.close()
    assert not bad, f"worlds differing from the oracle: {bad}"
    if tied:
        print(f"rerun of {len(tied)} worlds with 25-run chains", flush=True)
        b = twins_amd.Batch([worlds[i] for i in tied], chain_runs=25)
        b.run()
        for j, i in enumerate(tied):
            rows, ev = oracle(wargs[i], str(tmp_path / f"r{i:03d}.csv"), runs=25)
            est, done = b.fct(j)
            assert [(str(e), str(d)) for e, d in zip(est, done)] == rows and b.counters(j)[26] == ev, i
        b.close()
