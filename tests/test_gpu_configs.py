"""The benchmark configurations on the GPU against the oracle (no stored ns-3 output exists for
them): config 1 (FatTree k=4, lognormal flows) and config 2 (FatTree k=16 topology file, Pareto
flows, persistent connections, spray, FIN, a link failure) under each TCP; every flow's times and
the packet-event count must be identical."""
import os
import subprocess

import pytest

import util


def oracle(flows, args, out):
    r = subprocess.run([util.oracle_bin()] + args + [f"--flows={flows}", f"--out={out}"], capture_output=True, text=True, check=True)
    ev = int([l for l in r.stdout.splitlines() if l.startswith("packet-events")][0].split()[1])
    return util.read_cols(out, (4, 5)), ev


@pytest.mark.gpu
@pytest.mark.parametrize("name,tcp", [("config1", "dctcp"), ("config2", "newreno"), ("config2", "dctcp"), ("config2", "cubic")])
def test_config_matches_oracle(name, tcp, tmp_path):
    import twins_amd
    from twins_amd import workloads
    print(f"case {name} {tcp}", flush=True)
    cfg = workloads.write_configs(os.path.join(util.ROOT, "tests", "golden", "configs"))
    flows, args = cfg[name]
    args = [a for a in args if not a.startswith("--tcp=")] + [f"--tcp={tcp}"]
    ref, ev = oracle(flows, args, str(tmp_path / "o.csv"))
    b = twins_amd.Batch([twins_amd.World(args + [f"--flows={flows}"])])
    st = b.run()
    est, done = b.fct(0)
    assert [(str(e), str(d)) for e, d in zip(est, done)] == ref
    assert st.packet_events == ev
