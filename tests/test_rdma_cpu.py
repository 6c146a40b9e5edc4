"""The RDMA model's code on the CPU (tests/native/rdma_cpu: the slot kernel's steps run one entity
after the other, with the unchanged host code and command line) against the reference's RDMA jobs:
flow completions and counters identical, job by job. Also the command line's and the library's
error behaviour, and that the library refuses to run without a GPU."""
import os
import subprocess

import pytest

import rdma_util
import util


@pytest.fixture(scope="module")
def cpu_bin():
    subprocess.run(["make", "-C", os.path.dirname(rdma_util.CPU_BIN)], check=True, capture_output=True)
    return rdma_util.CPU_BIN


def run_cli(binary, topo, flows, kv, out, timeout=300):
    I = rdma_util.INPUTS
    return subprocess.run([binary, rdma_util.CONFIG, f"--topo={I}/{topo}", f"--flows={I}/{flows}", f"--out={out}"] + kv,
                          capture_output=True, text=True, timeout=timeout)


FAST = rdma_util.jobs(max_ref_s=1.0)


@pytest.mark.parametrize("job", FAST, ids=[j[0] for j in FAST])
def test_cpu_build_matches_reference_job(cpu_bin, job, tmp_path):
    name, topo, flows, kv, _, counters = job
    out = tmp_path / (name + ".fct")
    r = run_cli(cpu_bin, topo, flows, kv, out)
    assert r.returncode == 0, r.stderr
    assert out.read_text() == rdma_util.ref_fct(name)
    assert [l for l in r.stdout.splitlines() if l.startswith("drops ")] == [counters]


def test_cpu_build_batch_equals_single_runs(cpu_bin, tmp_path):
    """A job file of several worlds (the reference's --batch) gives each world's single-run results."""
    sel = FAST[:6]
    jf = tmp_path / "jobs.txt"
    jf.write_text("".join(f"{n} {t} {f} {' '.join(kv)}\n" for n, t, f, kv, _, _ in sel))
    r = subprocess.run([cpu_bin, rdma_util.CONFIG, f"--batch={jf}", f"--outdir={tmp_path}", f"--path={rdma_util.INPUTS}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    for n, *_ in sel:
        assert (tmp_path / (n + ".fct")).read_text() == rdma_util.ref_fct(n), n
    got = [l for l in r.stdout.splitlines() if l.startswith("drops ")]
    assert rdma_util.sum_counters(got) == rdma_util.sum_counters([j[5] for j in sel])


def test_serving_loop_on_cpu(tmp_path):
    """The serving loop (trd_server_*) with the kernel's steps on the CPU: four reference jobs in an
    arena that holds the two k=8 worlds only one at a time (admission waits for memory), every
    request's flows identical to the reference's."""
    d = os.path.dirname(rdma_util.CPU_BIN)
    subprocess.run(["make", "-C", d, "server_cpu"], check=True, capture_output=True)
    jobs = [("a", "ft4.topo", "inc1.csv", "CC_MODE=1,BUFFER_SIZE=1", "inc1_b1_cc1"),
            ("b", "fattree8.txt", "k8_incast16.csv", "CC_MODE=3,BUFFER_SIZE=2", "k8_incast16_b2_cc3"),
            ("c", "fattree8.txt", "k8_incast16.csv", "CC_MODE=1", "k8_incast16_cc1"),
            ("d", "ft4.topo", "inc1.csv", "CC_MODE=3,BUFFER_SIZE=1", "inc1_b1_cc3")]
    args = [os.path.join(d, "server_cpu"), rdma_util.CONFIG, rdma_util.INPUTS, str(tmp_path), "700"]
    for j in jobs:
        args += list(j[:4])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "most waiting 1" in r.stdout or "most waiting 2" in r.stdout, r.stdout
    for name, *_, ref in jobs:
        assert (tmp_path / (name + ".fct")).read_text() == rdma_util.ref_fct(ref), name


@pytest.mark.parametrize("kv,msg", [(["CC_MODE=2"], "only 1 (DCQCN), 3 (HPCC) and 8 (DCTCP)"), (["CC_MODE=3", "MULTI_RATE=1"], "MULTI_RATE 1"),
                                    (["L2_ACK_INTERVAL=0"], "L2_ACK_INTERVAL 0"), (["--bogus=1"], "unknown argument")])
def test_cli_errors(cpu_bin, kv, msg, tmp_path):
    r = run_cli(cpu_bin, "ft4.topo", "inc1.csv", kv, tmp_path / "x.fct")
    assert r.returncode == 1 and msg in r.stderr


@pytest.mark.parametrize("job", FAST, ids=[j[0] for j in FAST])
def test_cpu_serial_switch_path_matches_reference_job(cpu_bin, job, tmp_path):
    """Every switch on the serial path that switches of more than 32 ports take (switch_step, ports
    in device memory, lane 0 runs all events): the reference's jobs stay identical, counters too."""
    name, topo, flows, kv, _, counters = job
    out = tmp_path / (name + ".fct")
    I = rdma_util.INPUTS
    r = subprocess.run([cpu_bin, rdma_util.CONFIG, f"--topo={I}/{topo}", f"--flows={I}/{flows}", f"--out={out}"] + kv,
                       capture_output=True, text=True, timeout=300, env=dict(os.environ, TWINS_RDMA_CPU_SERIAL="1"))
    assert r.returncode == 0, r.stderr
    assert out.read_text() == rdma_util.ref_fct(name)
    assert [l for l in r.stdout.splitlines() if l.startswith("drops ")] == [counters]


@pytest.mark.parametrize("run", rdma_util.wide_runs(), ids=[r[0] for r in rdma_util.wide_runs()])
def test_cpu_wide_switch_world(cpu_bin, run, tmp_path):
    """Leaves of 44 ports (the serial path): the stored flow completions, every flow done, and no
    flow faster than its bytes at the line rate plus two link delays."""
    name, kv = run
    out = tmp_path / "w.fct"
    r = run_cli(cpu_bin, rdma_util.WIDE_TOPO, rdma_util.WIDE_FLOWS, kv, out)
    assert r.returncode == 0, r.stderr
    assert out.read_text() == open(os.path.join(rdma_util.RG, "ref", f"wide_ls2x40_{name}.fct")).read()
    sizes = [int(l.split(",")[2]) for l in open(os.path.join(rdma_util.INPUTS, rdma_util.WIDE_FLOWS)).read().splitlines()[1:]]
    fct = {int(l.split()[0]): int(l.split()[2]) for l in out.read_text().splitlines()}
    assert len(fct) == len(sizes) == 240
    assert all(fct[i] >= sizes[i] * 8 // 100 + 2000 for i in fct)


TX = rdma_util.tx_jobs()


@pytest.mark.parametrize("job", TX, ids=[j[0] for j in TX])
def test_cpu_build_txtrace_equals_reference(cpu_bin, job, tmp_path):
    """--txtrace (every transmission start in [from, until]) equals the reference's own trace file
    of the job, byte for byte (PFC PAUSE/RESUME, INT-carrying ACKs, ECN marks, graph topologies)."""
    name, topo, flows, kv, _, _ = job
    tx = tmp_path / "tx.txt"
    w0, w1 = rdma_util.TX_WINDOW
    r = run_cli(cpu_bin, topo, flows, kv + [f"--txtrace={tx}", f"--from={w0}", f"--until={w1}"], tmp_path / "x.fct")
    assert r.returncode == 0, r.stderr
    assert tx.read_bytes() == rdma_util.ref_txtrace(name)


def test_cli_txtrace_one_world_only(cpu_bin, tmp_path):
    jobs = tmp_path / "jobs.txt"
    jobs.write_text("a ft4.topo inc1.csv CC_MODE=1\n")
    r = subprocess.run([cpu_bin, rdma_util.CONFIG, f"--batch={jobs}", f"--outdir={tmp_path}", f"--path={rdma_util.INPUTS}",
                        f"--txtrace={tmp_path}/t.txt"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--txtrace traces one world" in r.stderr


def test_cli_sent_file(cpu_bin, tmp_path):
    """--sent: the reference records first SendCompletes only under the SimAI transmit rule, so for
    these worlds its file is empty (checked against its CPU build where that is built)."""
    sent = tmp_path / "x.sent"
    r = run_cli(cpu_bin, "ft4.topo", "inc1.csv", ["CC_MODE=1", "BUFFER_SIZE=1", f"--sent={sent}"], tmp_path / "x.fct")
    assert r.returncode == 0, r.stderr
    assert sent.exists() and sent.read_text() == ""
    ref = os.path.join(util.ROOT, "oracle", "_ref", "twins-rdma-cpu")
    if os.path.exists(ref):
        rs = tmp_path / "r.sent"
        rr = run_cli(ref, "ft4.topo", "inc1.csv", ["CC_MODE=1", "BUFFER_SIZE=1", f"--sent={rs}"], tmp_path / "r.fct")
        assert rr.returncode == 0 and rs.read_text() == sent.read_text()


def test_cli_refuses_what_is_not_modelled(cpu_bin, tmp_path):
    r = run_cli(cpu_bin, "ft4.topo", "inc1.csv", ["--astra=x"], tmp_path / "x.fct")
    assert r.returncode == 1 and "not modelled" in r.stderr
    dag = tmp_path / "x.dag"
    dag.write_text("flow 0 0 1 1000\nroot 0 0\n")
    r = subprocess.run([cpu_bin, rdma_util.CONFIG, f"--topo={rdma_util.INPUTS}/ft4.topo", f"--flows={dag}", f"--out={tmp_path}/y.fct"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "DAG flow files are not modelled" in r.stderr


def test_cli_missing_file(cpu_bin, tmp_path):
    r = run_cli(cpu_bin, "nonexistent.topo", "inc1.csv", [], tmp_path / "x.fct")
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_world_build_on_host_and_no_cpu_fallback():
    import twins_amd
    from twins_amd import rdma
    I = rdma_util.INPUTS
    w = rdma.World(rdma_util.CONFIG, f"{I}/ft4.topo", f"{I}/inc1.csv", ["CC_MODE=1"])
    assert (w.hosts, w.switches, w.flows) == (16, 20, 15)
    # maxRtt: 6 hops of 1 us there and back plus 6 MTU frames (1000 B at 100 Gb/s, 80 ns each);
    # maxBdp: its bytes at 100 Gb/s
    assert w.max_rtt == 2 * 6 * 1000 + 6 * 80 and w.max_bdp == w.max_rtt * 100 // 8
    with pytest.raises(twins_amd.TwError, match="CC_MODE 5"):
        rdma.World(rdma_util.CONFIG, f"{I}/ft4.topo", f"{I}/inc1.csv", ["CC_MODE=5"])
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(twins_amd.TwError, match="no GPU"):
        rdma.Batch([w])


def test_cli_lost_tails_rerun(cpu_bin, tmp_path):
    """Lost tails (src/rdma/rmain.cu:2326-2363): with rings of 2^10 nodes (--tail-auto=10) a tie of
    this PFC incast meets an overwritten tail node; the command prints the reference's warning,
    runs the worlds again with 2^26-node rings in the same process and ends with the reference's
    flow completions and counters."""
    name = "inc1_b1_cc1"
    out = tmp_path / "x.fct"
    r = run_cli(cpu_bin, "ft4.topo", "inc1.csv", ["CC_MODE=1", "BUFFER_SIZE=1", "--tail-auto=10"], out)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("drops ")]
    assert len(lines) == 2 and "on lost tails 0 " not in lines[0], r.stdout
    assert "warning: ties the chains did not decide exactly" in r.stderr and "a larger --tail helps" in r.stderr
    assert "lost tails: running again with --tail-auto=26" in r.stderr
    assert lines[1] == rdma_util.jobs()[[j[0] for j in rdma_util.jobs()].index(name)][5]
    assert out.read_text() == rdma_util.ref_fct(name)


def test_cli_tail_set_never_grows(cpu_bin, tmp_path):
    """--tail=N fixes the rings (the reference reruns only when --tail was not given): one run, the
    lost tails reported, no rerun."""
    r = run_cli(cpu_bin, "ft4.topo", "inc1.csv", ["CC_MODE=1", "BUFFER_SIZE=1", "--tail=10"], tmp_path / "x.fct")
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("drops ")]
    assert len(lines) == 1 and "on lost tails 0 " not in lines[0]
    assert "running again" not in r.stderr and "a larger --tail helps" in r.stderr
    r = run_cli(cpu_bin, "ft4.topo", "inc1.csv", ["CC_MODE=1", "--tail=3"], tmp_path / "y.fct")
    assert r.returncode == 1 and "from 4 to 30" in r.stderr


SERIAL = [j for j in rdma_util.jobs(max_ref_s=1.5) if "PFC" in j[0] or "b1" in j[0] or "incast" in j[0]][:8]


@pytest.mark.parametrize("job", SERIAL, ids=[j[0] for j in SERIAL])
def test_cli_serial_flag(cpu_bin, job, tmp_path):
    """--serial (src/rdma/rmain.cu:1090-1091, 1949): every switch slot in event order on one lane,
    never the per-port safe path; completions and counters stay the reference's."""
    name, topo, flows, kv, _, counters = job
    out = tmp_path / (name + ".fct")
    r = run_cli(cpu_bin, topo, flows, kv + ["--serial"], out)
    assert r.returncode == 0, r.stderr
    assert out.read_text() == rdma_util.ref_fct(name)
    assert [l for l in r.stdout.splitlines() if l.startswith("drops ")][-1] == counters
