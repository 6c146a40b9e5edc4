"""The serving loop (tw_server_*): requests submitted while others run on the device give each
world's own results. Worlds of all three queue sizes (FatTree k=8, k=16, and a k=32 world) are
submitted at once and in a second wave while the first runs; every flow must equal ns-3's stored
output (or, for the generated k=32 world, the oracle's), as the plain path's do."""
import os
import subprocess
import time

import pytest

import util


def _golden(names):
    by = {c[0]: c for c in util.cases()}
    out = []
    for n in names:
        _, fl, ref, args = by[n]
        out.append(([a.replace("topo/", util.GOLDEN + "/topo/") for a in args] + [f"--flows={util.GOLDEN}/{fl}"],
                    util.read_cols(os.path.join(util.GOLDEN, ref), (4, 5))))
    return out


@pytest.mark.gpu
def test_server_results_equal_reference(tmp_path):
    import twins_amd
    from twins_amd import workloads
    first = _golden(["k8_paper-dctcp-base", "k8_incast16-cubic-base", "k16_paper-dctcp-base"])
    second = _golden(["k8_a2a30-newreno-base", "k16_a2a30-dctcp-base", "k8_paper-newreno-spray"])
    # a k=32 world (queue of the largest worlds), checked against the oracle
    reqs = workloads.config5_requests(str(tmp_path / "c5"), n_requests=1, seed=55, ks=(32,), max_flows=2000, fail_share=1.0)
    big_args = reqs[0][1]
    o = str(tmp_path / "o.csv")
    subprocess.run([util.oracle_bin()] + big_args + [f"--out={o}"], check=True, capture_output=True)
    big = (big_args, util.read_cols(o, (4, 5)))
    print("server: 3 + big, then 3 more while they run", flush=True)
    srv = twins_amd.Server(arena_bytes=24 << 30, idle_s=1.0)
    jobs = {}
    for args, exp in first + [big]:
        jobs[srv.submit(twins_amd.World(args))] = exp
    time.sleep(0.05)
    for args, exp in second:
        jobs[srv.submit(twins_amd.World(args))] = exp
    t0 = time.time()
    left = set(jobs)
    while left and time.time() - t0 < 90:
        st = srv.poll()
        for rid in list(left):
            if srv.state(rid) in (twins_amd.REQ_DONE, twins_amd.REQ_FAILED):
                est, done, ev = srv.result(rid)
                got = [(str(e), str(d)) for e, d in zip(est, done)]
                assert got == jobs[rid], f"request {rid}"
                left.discard(rid)
        time.sleep(0.01)
    srv.close()
    assert not left, f"requests {sorted(left)} not finished in 90 s"
