"""`twins --ecmp_from=F --hash_check=1` (src/main.cu:723-768; the reference's validation/check_hashes.sh):
the ECMP hashes the product computes for every connection (ns-3's Hash32 over the 5-tuple, with the
ephemeral ports ns-3 assigns) against the hash_data / hash_ack columns of every stored ns-3 output,
without simulating (host code only: no GPU needed). A hash changed in one row must be reported."""
import os
import subprocess

import pytest

import util

CASES = util.cases()


def hash_check(args, flows, ref):
    return subprocess.run([util.TWINS] + args + [f"--flows={flows}", f"--ecmp_from={ref}", "--hash_check=1"], cwd=util.GOLDEN,
                          capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module")
def twins_bin():
    if not os.path.exists(util.TWINS):
        pytest.fail("twins_amd/bin/twins is not built (python __graft_entry__.py build)")
    return util.TWINS


@pytest.mark.parametrize("name,flows,ref,args", CASES, ids=[c[0] for c in CASES])
def test_hashes_equal_ns3(twins_bin, name, flows, ref, args):
    r = hash_check(args, flows, ref)
    n = len(util.read_cols(os.path.join(util.GOLDEN, ref), (6,)))
    assert r.returncode == 0, r.stdout + r.stderr
    line = r.stdout.strip().splitlines()[-1]
    a, b = int(line.split()[2]), int(line.split()[4])
    assert line.startswith("ECMP hashes: ") and a == b and 0 < b <= n


def test_hash_check_reports_a_wrong_hash(twins_bin, tmp_path):
    name, flows, ref, args = next(c for c in CASES if c[0] == "k8_r30-dctcp-sprayfin")
    rows = open(os.path.join(util.GOLDEN, ref)).read().splitlines()
    v = rows[1].split(",")
    v[6] = str((int(v[6]) + 1) % (1 << 32))
    rows[1] = ",".join(v)
    bad = tmp_path / "bad.csv"
    bad.write_text("\n".join(rows) + "\n")
    r = hash_check(args, flows, str(bad))
    assert r.returncode == 1 and "ECMP hashes: 167 of 168 connections as ns-3's" in r.stdout
