"""The GPU path against sequential ns-3's stored outputs, through the drop-in `twins` command line
(the C ABI underneath): every flow's establishment and completion time must be identical, as the
reference's validation/check_exact.sh demands. Undecided ties are rerun with 25-run chains."""
import os

import pytest

import util

CASES = util.cases()


def run_gpu(args, flows, out, runs=9):
    return util.run(util.TWINS, args + [f"--flows={flows}", f"--chain_runs={runs}"], out)


@pytest.mark.gpu
@pytest.mark.parametrize("name,flows,ref,args", CASES, ids=[c[0] for c in CASES])
def test_gpu_matches_ns3(name, flows, ref, args, tmp_path):
    print(f"case {name}", flush=True)
    out = str(tmp_path / "g.csv")
    log = run_gpu(args, flows, out)
    same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    if same != n and util.undecided(log):
        log = run_gpu(args, flows, out, runs=25)
        same, n = util.same_times(os.path.join(util.GOLDEN, ref), out)
    assert same == n, f"{name}: {same}/{n} flows identical to ns-3\n{log[-1500:]}"


@pytest.mark.gpu
def test_gpu_batch_equals_standalone(tmp_path):
    """Heterogeneous worlds in one batch (one persistent launch) give each world's own ns-3
    output: batching changes nothing (the reference's check_sweeps.sh property)."""
    import twins_amd
    pick = [c for c in CASES if c[0].split("-")[2] in ("base", "spray", "fin") and not c[0].startswith("k16")][:24]
    worlds = [twins_amd.World([a.replace("topo/", util.GOLDEN + "/topo/") for a in args] + [f"--flows={util.GOLDEN}/{fl}"])
              for _, fl, _, args in pick]
    print(f"batch of {len(worlds)} worlds", flush=True)
    b = twins_amd.Batch(worlds)
    st = b.run()
    assert st.worlds_done == len(worlds)
    tied = []
    for i, (name, fl, ref, args) in enumerate(pick):
        est, done = b.fct(i)
        exp = util.read_cols(os.path.join(util.GOLDEN, ref), (4, 5))
        got = [(str(e), str(d)) for e, d in zip(est, done)]
        if b.counters(i)[25]:
            tied.append(i)  # undecided ties with 9-run chains: run again below with 25
            continue
        assert got == exp, name
    b.close()
    if not tied:
        return
    # the worlds whose chains left ties undecided, again in one batch of 25-run chains (the
    # reference's twins128 rerun): each must now equal its ns-3 output
    print(f"rerun of {len(tied)} worlds with 25-run chains: {[pick[i][0] for i in tied]}", flush=True)
    b = twins_amd.Batch([worlds[i] for i in tied], chain_runs=25)
    assert b.run().worlds_done == len(tied)
    for j, i in enumerate(tied):
        name, fl, ref, args = pick[i]
        est, done = b.fct(j)
        exp = util.read_cols(os.path.join(util.GOLDEN, ref), (4, 5))
        assert b.counters(j)[25] == 0, name
        assert [(str(e), str(d)) for e, d in zip(est, done)] == exp, name
    b.close()
