"""The slot kernel's model code (twins_amd/csrc/tcp_model.h, the functions the gfx950 kernel runs)
compiled for the CPU under AddressSanitizer by tests/native/cpu_harness, in the kernel's order over
the same batch layout: every memory access is bounds-checked, and the flow times must equal the
stored ns-3 outputs. This checks the kernel code itself without a GPU."""
import os
import subprocess

import pytest

import util

HARNESS = os.path.join(util.ROOT, "tests", "native", "cpu_harness")
BIN = os.path.join(HARNESS, "twins_cpucheck")
CASES = ["k8_paper-dctcp-base", "k8_paper-cubic-jit", "k8_r30-newreno-sprayfin", "ls_incast16-newreno-base", "ft8f_a2a30-dctcp-base"]


@pytest.fixture(scope="module")
def harness():
    subprocess.run(["make", "-C", HARNESS], check=True, capture_output=True, timeout=900)
    return BIN


@pytest.mark.parametrize("name", CASES)
def test_kernel_code_on_cpu_matches_ns3(harness, name, tmp_path):
    by_name = {c[0]: c for c in util.cases()}
    if name not in by_name:
        pytest.skip(f"{name} not in the manifest")
    _, fl, ref, args = by_name[name]
    out = str(tmp_path / "fct.csv")
    r = subprocess.run([harness] + args + [f"--flows={fl}", f"--out={out}"], cwd=util.GOLDEN, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    assert same == n, f"{name}: {same}/{n} flows identical"


def test_kernel_code_on_cpu_overflow_regime_matches_oracle(harness, tmp_path):
    """16 hosts with 56,000 flows: active lists overflow (303 connections refused, 30 flows never
    finish). The reference's results in that regime depend on when finished connections leave the
    lists (at the end of every slot a host runs); the kernel code must do the same, flow for flow
    and in its counters."""
    from twins_amd import workloads
    flows = tmp_path / "ovf.csv"
    flows.write_text(workloads.pareto_flows(5000, 56000, 16))
    args = ["--k=4", f"--flows={flows}", "--tcp=cubic"]
    o, h = str(tmp_path / "o.csv"), str(tmp_path / "h.csv")
    ro = subprocess.run([util.oracle_bin()] + args + [f"--out={o}"], capture_output=True, text=True, check=True, timeout=600)
    rh = subprocess.run([harness] + args + [f"--out={h}"], capture_output=True, text=True, timeout=600)
    assert rh.returncode == 0, rh.stderr[-3000:]
    ovf = lambda out: int(out.split("active-ovf ")[1].split()[0])
    assert ovf(ro.stdout) == ovf(rh.stdout) > 0
    same, n = util.same_times(o, h)
    assert same == n


QT_CASES = ["k8_r30-newreno-sprayfin", "ls_incast16-newreno-base", "ft8f_a2a30-dctcp-base"]


@pytest.mark.parametrize("name", QT_CASES)
def test_kernel_code_qtrace_equals_oracle(harness, name, tmp_path):
    """--qtrace (every enqueue in [from, until), src/main.cu:1077-1100) from the kernel code is the
    oracle's, byte for byte, over a window of the run (spray, leaf-spine, failures)."""
    by_name = {c[0]: c for c in util.cases()}
    _, fl, ref, args = by_name[name]
    win = ["--qtrace_from=200000000", "--qtrace_until=400000000"]
    o, h = tmp_path / "o.txt", tmp_path / "h.txt"
    subprocess.run([util.oracle_bin()] + args + [f"--flows={fl}", f"--out={tmp_path / 'o.csv'}", f"--qtrace={o}"] + win, cwd=util.GOLDEN,
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([harness] + args + [f"--flows={fl}", f"--out={tmp_path / 'h.csv'}", f"--qtrace={h}"] + win, cwd=util.GOLDEN,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    a, b = o.read_bytes(), h.read_bytes()
    assert len(a) > 10000 and a == b


def test_noalloc_entry_on_cpu(harness, tmp_path):
    """tw_batch_run_noalloc through `twins --noalloc=1`: the launch's record, barrier words, queue
    counters and switch sort space moved into the static scratch array (host memory here), every
    access bounds-checked; the flows equal ns-3's. A launch larger than the array is refused."""
    by_name = {c[0]: c for c in util.cases()}
    _, fl, ref, args = by_name["k8_r30-newreno-sprayfin"]
    out = str(tmp_path / "fct.csv")
    r = subprocess.run([harness] + args + [f"--flows={fl}", f"--out={out}", "--noalloc=1", "--blocks=8"], cwd=util.GOLDEN,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    assert same == n
    r = subprocess.run([harness] + args + [f"--flows={fl}", f"--out={out}", "--noalloc=1", "--blocks=40", "--team_blocks=40"],
                       cwd=util.GOLDEN,
                       capture_output=True, text=True, timeout=600, env={**os.environ, "ASAN_OPTIONS": "detect_leaks=0"})
    assert r.returncode == 1 and "static scratch" in r.stderr
