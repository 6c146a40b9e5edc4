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
