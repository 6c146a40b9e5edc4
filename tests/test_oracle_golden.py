"""The oracle (CPU restatement of the model) against sequential ns-3's stored outputs: every flow's
establishment and completion time, and the ECMP hash columns. A case whose recorded histories
leave a tie undecided is rerun with 25-run chains, as the reference does with twins128."""
import os

import pytest

import util

SMALL = [c for c in util.cases() if not c[0].startswith("k16")]


@pytest.mark.parametrize("name,flows,ref,args", SMALL, ids=[c[0] for c in SMALL])
def test_oracle_matches_ns3(name, flows, ref, args, oracle_bin, tmp_path):
    out = str(tmp_path / "o.csv")
    log = util.run(oracle_bin, args + [f"--flows={flows}", "--hash_cols=1"], out)
    same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    if same != n and util.undecided(log):
        log = util.run(oracle_bin, args + [f"--flows={flows}", "--hash_cols=1", "--chain_runs=25"], out)
        same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    assert same == n, f"{name}: {same}/{n} flows identical to ns-3"
    # the ns-3 ECMP hashes of every connection (what --hash_check verifies in the reference)
    assert util.read_cols(os.path.join(util.GOLDEN, ref), (6, 7)) == util.read_cols(out, (6, 7))


def test_oracle_plants_are_caught(oracle_bin, tmp_path):
    """A planted mistake in the run (a RED threshold off by one) must break parity: the golden
    comparison is sharp enough to notice a one-packet change in one queue parameter."""
    name, flows, ref, args = next(c for c in SMALL if c[0] == "k8_incast16-dctcp-base")
    out = str(tmp_path / "o.csv")
    util.run(oracle_bin, args + [f"--flows={flows}", "--ecnK=31"], out)
    same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    assert same < n
