"""The queue trace (`--qtrace`, the reference's record of every enqueue: src/main.cu:1077-1100,
src/step.cuh:266-279) from the GPU against the oracle's, byte for byte: stored golden cases through
the `twins` command line over a window of their run, and config 2 (FatTree k=16, spray, FIN, a link
failure) over its whole run. A batch of worlds traces all of them: one record per packet-event."""
import os
import subprocess

import pytest

import util

QT_CASES = ["k8_r30-newreno-sprayfin", "ls_incast16-newreno-base", "ft8f_a2a30-dctcp-base"]
WINDOW = ["--qtrace_from=200000000", "--qtrace_until=400000000"]


def oracle_trace(args, flows, tmp_path, extra=()):
    qt = tmp_path / "o.txt"
    r = subprocess.run([util.oracle_bin()] + args + [f"--flows={flows}", f"--out={tmp_path / 'o.csv'}", f"--qtrace={qt}"] + list(extra),
                       cwd=util.GOLDEN, capture_output=True, text=True, check=True, timeout=600)
    return qt.read_bytes(), util.packet_events(r.stdout)


@pytest.mark.gpu
@pytest.mark.parametrize("name", QT_CASES)
def test_gpu_qtrace_window_equals_oracle(name, tmp_path):
    print(f"qtrace case {name}", flush=True)
    by_name = {c[0]: c for c in util.cases()}
    _, fl, _, args = by_name[name]
    want, _ = oracle_trace(args, fl, tmp_path, WINDOW)
    qt = tmp_path / "g.txt"
    log = util.run(util.TWINS, args + [f"--flows={fl}", f"--qtrace={qt}"] + WINDOW, str(tmp_path / "g.csv"))
    got = qt.read_bytes()
    lines = got.count(b"\n")
    assert f"qtrace: {lines} records" in log
    assert len(want) > 10000 and got == want


@pytest.mark.gpu
def test_gpu_qtrace_config2(tmp_path):
    import twins_amd
    from twins_amd import workloads
    cfg = workloads.write_configs(os.path.join(util.ROOT, "tests", "golden", "configs"))
    flows, args = cfg["config2"]
    base = [a for a in args if not a.startswith("--tcp=")]
    whole = ["--qtrace_until=100000000000000"]
    print("qtrace config2 dctcp", flush=True)
    want, ev = oracle_trace(base + ["--tcp=dctcp"], flows, tmp_path, whole)
    b = twins_amd.Batch([twins_amd.World(base + ["--tcp=dctcp", f"--flows={flows}"])])
    b.qtrace_enable(0, 100000000000000)
    st = b.run()
    assert st.packet_events == ev
    n = b.qtrace_write(tmp_path / "g.txt")
    assert n == ev and (tmp_path / "g.txt").read_bytes() == want
    # three worlds in one batch: every world's enqueues, one record per packet-event
    print("qtrace config2 batch of 3", flush=True)
    b3 = twins_amd.Batch([twins_amd.World(base + [f"--tcp={t}", f"--flows={flows}"]) for t in ("newreno", "dctcp", "cubic")])
    b3.qtrace_enable(0, 100000000000000)
    st3 = b3.run()
    n3 = b3.qtrace_write(tmp_path / "g3.txt")
    assert n3 == st3.packet_events
    keys = [tuple(map(int, l.split()[:3])) for l in open(tmp_path / "g3.txt")]
    assert len(keys) == n3 and keys == sorted(keys)
