"""The RDMA model on the GPU against the reference's RDMA jobs (tests/golden/rdma): every world's
flow completions identical to the reference's, and the batch's counters equal to the sum of the
jobs' counters. Worlds run side by side in one batch (the reference's --batch), through the C ABI."""
import os
import subprocess
import time

import pytest

import rdma_util

pytestmark = pytest.mark.gpu


def run_batch(sel, label):
    from twins_amd import rdma
    I = rdma_util.INPUTS
    print(f"rdma batch {label}: {len(sel)} worlds", flush=True)
    worlds = [rdma.World(rdma_util.CONFIG, f"{I}/{t}", f"{I}/{f}", kv) for _, t, f, kv, _, _ in sel]
    b = rdma.Batch(worlds)
    t = time.time()
    st = b.run()
    print(f"  {time.time() - t:.1f} s wall, {st.kernel_s:.1f} s kernel, {st.launches} launches", flush=True)
    bad = [n for i, (n, *_) in enumerate(sel) if b.fct_text(i) != rdma_util.ref_fct(n)]
    assert not bad, f"worlds differing from the reference: {bad}"
    got, want = rdma_util.sum_counters([b.summary_line()]), rdma_util.sum_counters([j[5] for j in sel])
    # every counter equal, undecided ties and ties on lost tails included (each world has a tail ring
    # of its own, so worlds running side by side do not overwrite each other's chain tails)
    print(f"  undecided ties {got[5]} (reference {want[5]}), on lost tails {got[6]} (reference {want[6]})", flush=True)
    assert got == want


def test_gpu_reference_jobs_short():
    """The 80 jobs the reference's CPU build runs in up to 3 s, in one batch."""
    run_batch(rdma_util.jobs(max_ref_s=3.0), "short")


def test_gpu_reference_jobs_medium():
    """The 36 jobs the reference's CPU build runs in 3-50 s (FatTree k=16, leaf-spine 256 hosts,
    PFC sweeps), in one batch."""
    run_batch(rdma_util.jobs(min_ref_s=3.0, max_ref_s=50.0), "medium")


def test_gpu_reference_jobs_long():
    """Two of the 4 jobs the reference's CPU build runs in 70-410 s: the DCTCP-RDMA permutations on
    FatTree k=16 with short flows (about 72 s each there), in one batch. (The other two, k=8 with
    long flows over 476,000 slots, are identical too but take 142 s on the GPU: too long for this
    suite.)"""
    run_batch([j for j in rdma_util.jobs(min_ref_s=50.0) if j[0].startswith("dons_k16")], "long")


@pytest.mark.parametrize("name", ["hpcc", "dcqcn"])
def test_gpu_config3(name):
    """Config 3 (512-host leaf-spine, all-to-all + 160-to-1 incast) against the reference's output."""
    from twins_amd import rdma
    import util
    runs = {l.split("|")[0]: l.rstrip("\n").split("|", 2) for l in open(os.path.join(rdma_util.RG, "config3_manifest.txt")) if not l.startswith("#")}
    _, kv, counters = runs[name]
    cfgd = os.path.join(util.GOLDEN, "configs")
    print(f"rdma config 3 {name}", flush=True)
    w = rdma.World(rdma_util.CONFIG, os.path.join(cfgd, "ls32.txt"), os.path.join(cfgd, "config3_flows.csv"), kv.split())
    b = rdma.Batch([w], q3cap=16384)
    b.run()
    assert b.fct_text(0) == rdma_util.ref_fct(f"config3_{name}")
    assert b.summary_line() == counters


def test_gpu_cli_one_job(tmp_path):
    """The `twins-rdma` command line on the GPU: the reference's output file and summary lines."""
    import util
    I = rdma_util.INPUTS
    out = tmp_path / "x.fct"
    r = subprocess.run([os.path.join(util.ROOT, "twins_amd", "bin", "twins-rdma"), rdma_util.CONFIG, f"--topo={I}/fattree8.txt",
                        f"--flows={I}/k8_incast16.csv", f"--out={out}", "CC_MODE=3", "BUFFER_SIZE=2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_text() == rdma_util.ref_fct("k8_incast16_b2_cc3")
    job = [j for j in rdma_util.jobs() if j[0] == "k8_incast16_b2_cc3"][0]
    assert [l for l in r.stdout.splitlines() if l.startswith("drops ")] == [job[5]]
    assert "flows done 160 of 160" in r.stdout


def test_gpu_txtrace():
    """--txtrace on the GPU: the command line's trace of each fixture job equals the reference's
    own trace file; the same jobs as one batch give each world's trace in its own numbering; a whole
    run's trace has one record per packet-event (68,070 in the reference's run of inc1_b1_cc1)."""
    import tempfile
    import util
    from twins_amd import rdma
    I = rdma_util.INPUTS
    w0, w1 = rdma_util.TX_WINDOW
    sel = rdma_util.tx_jobs()
    with tempfile.TemporaryDirectory() as d:
        for name, topo, flows, kv, _, _ in sel:
            print(f"txtrace cli {name}", flush=True)
            tx = os.path.join(d, name + ".txt")
            r = subprocess.run([os.path.join(util.ROOT, "twins_amd", "bin", "twins-rdma"), rdma_util.CONFIG, f"--topo={I}/{topo}",
                                f"--flows={I}/{flows}", f"--out={d}/x.fct", f"--txtrace={tx}", f"--from={w0}", f"--until={w1}"] + kv,
                               capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            assert open(tx, "rb").read() == rdma_util.ref_txtrace(name), name
        print(f"txtrace batch of {len(sel)}", flush=True)
        b = rdma.Batch([rdma.World(rdma_util.CONFIG, f"{I}/{t}", f"{I}/{f}", kv) for _, t, f, kv, _, _ in sel])
        b.txtrace_enable(w0, w1)
        b.run()
        for i, (name, *_) in enumerate(sel):
            p = os.path.join(d, f"b{i}.txt")
            b.txtrace_write(i, p)
            assert open(p, "rb").read() == rdma_util.ref_txtrace(name), name
        print("txtrace whole run inc1_b1_cc1", flush=True)
        w = rdma.World(rdma_util.CONFIG, f"{I}/ft4.topo", f"{I}/inc1.csv", ["CC_MODE=1", "BUFFER_SIZE=1"])
        b1 = rdma.Batch([w])
        b1.txtrace_enable()
        b1.run()
        n, seen = b1.txtrace_write(0, os.path.join(d, "all.txt"))
        assert n == seen == b1.counters()[38] == 68070


def test_gpu_wide_switches():
    """A world whose leaves have 44 ports (more than a switch workspace's 32: the serial path in
    device memory) under DCQCN, HPCC, DCTCP and with PFC, as one batch: every run's flow completions
    equal the stored ones (make_wide.py)."""
    from twins_amd import rdma
    I = rdma_util.INPUTS
    runs = rdma_util.wide_runs()
    print(f"wide switches: {len(runs)} worlds", flush=True)
    b = rdma.Batch([rdma.World(rdma_util.CONFIG, f"{I}/{rdma_util.WIDE_TOPO}", f"{I}/{rdma_util.WIDE_FLOWS}", kv) for _, kv in runs])
    st = b.run()
    assert st.flows_done == 240 * len(runs)
    for i, (name, _) in enumerate(runs):
        assert b.fct_text(i) == open(os.path.join(rdma_util.RG, "ref", f"wide_ls2x40_{name}.fct")).read(), name


def test_gpu_rdma_serving_loop():
    """The RDMA serving loop: reference jobs of both size classes submitted one after another while
    earlier ones run; every request's flows identical to the reference's, whichever requests
    shared its launches."""
    import time
    import twins_amd
    from twins_amd import rdma
    I = rdma_util.INPUTS
    sel = [j for j in rdma_util.jobs(max_ref_s=3.0) if j[1] in ("ft4.topo", "fattree8.txt")][:10]
    print(f"rdma serving loop: {len(sel)} requests", flush=True)
    srv = rdma.Server()
    ids = []
    for n, t, f, kv, _, _ in sel:
        w = rdma.World(rdma_util.CONFIG, f"{I}/{t}", f"{I}/{f}", kv)
        ids.append((n, srv.submit(w)))
        srv.poll()
    t0 = time.time()
    done = {}
    while len(done) < len(ids) and time.time() - t0 < 100:
        srv.poll()
        for n, i in ids:
            if n not in done and srv.state(i) in (twins_amd.REQ_DONE, twins_amd.REQ_FAILED):
                assert srv.state(i) == twins_amd.REQ_DONE, n
                start, fct, ev = srv.result(i)
                done[n] = "".join(f"{j} {start[j]} {fct[j]}\n" for j in range(len(fct)) if fct[j] >= 0)
        time.sleep(0.001)
    srv.close()
    assert len(done) == len(ids)
    bad = [n for n, _ in ids if done[n] != rdma_util.ref_fct(n)]
    assert not bad, bad


def test_gpu_lost_tails_rerun():
    """Rings of 2^4 tail nodes per world (one reservation chunk, so every chunk an entity takes
    overwrites the nodes before it): ties of these incasts meet overwritten tail nodes (counter 9;
    the CPU build of the same code loses 13, 316 and 105 of them). run_exact() runs the worlds
    again from the start on 2^26-node rings, as the reference's twins-rdma reruns itself, and
    every world ends with the reference's completions and counters."""
    from twins_amd import rdma
    I = rdma_util.INPUTS
    names = ["inc1_b1_cc1", "jf32_incast_b1_cc1", "k8_perm_lowk_cc8"]
    by = {j[0]: j for j in rdma_util.jobs()}
    print(f"rdma lost tails: {names} on 2^4-node rings", flush=True)
    ws = [rdma.World(rdma_util.CONFIG, f"{I}/{by[n][1]}", f"{I}/{by[n][2]}", by[n][3]) for n in names]
    b = rdma.Batch(ws, tail_log2=4, tail_exact=True)
    first = b.run()
    lost = b.counters()[rdma.EV_LOST]
    assert lost > 0, "no tie met an overwritten node on 16-node rings"
    b.close()
    b = rdma.Batch(ws, tail_log2=4, tail_exact=True)
    st, reruns = b.run_exact()
    assert reruns >= 1 and b.tail_log2() >= 26
    assert b.counters()[rdma.EV_LOST] == 0
    for i, n in enumerate(names):
        assert b.fct_text(i) == rdma_util.ref_fct(n), n
    assert rdma_util.sum_counters([b.summary_line()]) == rdma_util.sum_counters([by[n][5] for n in names])
    b.close()


def test_gpu_serial_flag():
    """The reference's --serial on the GPU: every switch slot on one lane in event order (the safe
    per-port path off); 20 short reference jobs in one batch stay identical, counters included."""
    from twins_amd import rdma
    I = rdma_util.INPUTS
    sel = rdma_util.jobs(max_ref_s=1.0)[:20]
    print(f"rdma serial batch: {len(sel)} worlds", flush=True)
    b = rdma.Batch([rdma.World(rdma_util.CONFIG, f"{I}/{t}", f"{I}/{f}", kv) for _, t, f, kv, _, _ in sel], serial=True)
    b.run()
    bad = [n for i, (n, *_) in enumerate(sel) if b.fct_text(i) != rdma_util.ref_fct(n)]
    assert not bad, bad
    assert rdma_util.sum_counters([b.summary_line()]) == rdma_util.sum_counters([j[5] for j in sel])
    b.close()
