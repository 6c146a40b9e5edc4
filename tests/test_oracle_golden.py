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


def test_oracle_qtrace_invariants(oracle_bin, tmp_path):
    """The queue trace's numbering and fields, pinned by what the topology and TCP fix rather than by
    a second copy of the writer: k=4 FatTree (16 hosts; edge switch e = node 16 + e serves hosts 2e and
    2e + 1 on devices 1 and 2), 64 one-message connections, the whole run traced."""
    flows = os.path.join(util.GOLDEN, "configs", "config1_flows.csv")
    out, qt = str(tmp_path / "o.csv"), str(tmp_path / "q.txt")
    log = util.run(oracle_bin, ["--k=4", f"--flows={flows}", f"--qtrace={qt}", "--qtrace_until=1000000000000"], out)
    q = util.read_qtrace(qt)
    H = 16
    # every enqueued frame starts on its link before the run ends (all flows complete)
    assert len(q) == util.packet_events(log) > 10000
    assert q == sorted(q, key=lambda r: (r[0], r[1], r[2]))
    first_at = {}
    for ts, node, dev, src, dst, kind, sa, qlen in q:
        assert kind in ("D", "A", "S", "SA") and 1 <= qlen <= 200 and src != dst
        if node < H:
            assert dev == 1 and src == node  # a host's NIC carries only its own frames
        else:
            assert 1 <= dev <= 4 and node < H + 20
            if node < H + 8 and dev <= 2:  # an edge switch's down port leads to its destination host
                assert dst == 2 * (node - H) + dev - 1
        first_at.setdefault((node, dev), qlen)
    assert set(first_at.values()) == {1}  # every queue starts empty
    rows = [tuple(map(int, l.split(","))) for l in open(flows).read().split("\n")[1:] if l]
    pairs = {}
    for s, d, size, _ in rows:
        pairs.setdefault((s, d), []).append(size)
    for (s, d), sizes in pairs.items():
        if (d, s) in pairs:
            continue  # frames of both directions' connections share the pair
        mine = [r for r in q if (r[3], r[4]) == (s, d)]
        assert mine[0][1:3] == (s, 1) and mine[0][5] == "S" and mine[0][6] == 0  # the SYN leaves first
        back = [r for r in q if (r[3], r[4]) == (d, s)]
        assert back[0][1:3] == (d, 1) and back[0][5] == "SA"
        if len(sizes) == 1:
            seqs = {r[6] for r in mine if r[5] == "D" and r[1] == s}
            assert seqs >= set(range(0, sizes[0], 1460))  # relative sequence numbers of every segment
            acks = [r[6] for r in back if r[5] == "A" and r[1] == d]
            assert max(acks) == sizes[0]  # the last ACK covers every byte (numbers relative to the first)
