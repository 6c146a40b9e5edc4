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


@pytest.mark.gpu
def test_serving_mix_every_request(tmp_path):
    """Config 5 in small: 48 requests (FatTree k 4/8/16, up to 6,000 flows, a quarter with a link
    failure; 30% of them RDMA worlds under DCQCN, HPCC or DCTCP) arriving Poisson at 100/s, served
    by the TCP and RDMA serving loops side by side with the GPU split as the bench splits it. Every
    request is checked: TCP worlds flow by flow against the oracle, RDMA worlds against the
    reference's own RDMA model built for the CPU (oracle/_ref)."""
    import concurrent.futures as cf
    import twins_amd
    from twins_amd import rdma, workloads
    ref_bin = os.path.join(util.ROOT, "oracle", "_ref", "twins-rdma-cpu")
    if not os.path.exists(ref_bin):
        pytest.skip("oracle/_ref is not built (it needs the reference's sources at build time)")
    cfg = os.path.join(util.GOLDEN, "rdma", "inputs", "hpcc.cfg")
    reqs = workloads.config5_requests(str(tmp_path / "c5"), n_requests=48, seed=7, rate_per_s=100.0, ks=(4, 8, 16),
                                      max_flows=6000, rdma_share=0.3)
    is_rdma = [isinstance(r[1], dict) for r in reqs]
    assert 0 < sum(is_rdma) < len(reqs)
    worlds = [rdma.World(cfg, r[1]["topo"], r[1]["flows"], r[1]["overrides"]) if is_rdma[i] else twins_amd.World(r[1])
              for i, r in enumerate(reqs)]

    def expected(i):
        out = str(tmp_path / f"e{i:03d}")
        if is_rdma[i]:
            r = reqs[i][1]
            subprocess.run([ref_bin, cfg, f"--topo={r['topo']}", f"--flows={r['flows']}", f"--out={out}", "--gwin=0"] + r["overrides"],
                           check=True, capture_output=True, timeout=900)
            return open(out).read()
        subprocess.run([util.oracle_bin()] + reqs[i][1] + [f"--out={out}"], check=True, capture_output=True, timeout=900)
        return util.read_cols(out, (4, 5))
    # the references, computed on the host's cores while the GPU serves
    pool = cf.ThreadPoolExecutor(min(16, os.cpu_count() or 1))
    want = {i: pool.submit(expected, i) for i in range(len(reqs))}
    resident = twins_amd.launch_capacity(9, 256)
    rblocks = 64
    srv = twins_amd.Server(idle_s=2.0, small_teams=resident - resident // 8 - rblocks - 16 * 4 - 4 * 8)
    rsrv = rdma.Server(max_blocks=rblocks)
    print(f"serving mix: {len(reqs)} requests ({sum(is_rdma)} RDMA)", flush=True)
    sid, got = {}, {}
    t0 = time.perf_counter()
    nxt = 0
    while len(got) < len(reqs) and time.perf_counter() - t0 < 120:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] <= now:
            sid[nxt] = (rsrv if is_rdma[nxt] else srv).submit(worlds[nxt])
            nxt += 1
        srv.poll()
        rsrv.poll()
        for i in list(sid):
            if i in got:
                continue
            s = rsrv if is_rdma[i] else srv
            st = s.state(sid[i])
            if st not in (twins_amd.REQ_DONE, twins_amd.REQ_FAILED):
                continue
            assert st == twins_amd.REQ_DONE, f"request {i} failed"
            if is_rdma[i]:
                start, fct, ev = s.result(sid[i])
                assert ev[rdma.EV_LOST] == 0, f"request {i}: ties on lost tails after the server's reruns"
                got[i] = "".join(f"{j} {start[j]} {fct[j]}\n" for j in range(len(fct)) if fct[j] >= 0)
            else:
                est, done, ev = s.result(sid[i])
                got[i] = [(str(e), str(d)) for e, d in zip(est, done)]
        time.sleep(0.001)
    srv.close()
    rsrv.close()
    print(f"served {len(got)} of {len(reqs)} in {time.perf_counter() - t0:.1f} s", flush=True)
    assert len(got) == len(reqs), f"requests not finished in 120 s: {sorted(set(range(len(reqs))) - set(got))}"
    bad = [i for i in range(len(reqs)) if got[i] != want[i].result()]
    pool.shutdown()
    assert not bad, f"requests differing from the reference: {bad}"
