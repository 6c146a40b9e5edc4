"""tw_batch_run_noalloc: the same batch run with the launch's scratch in the library's static device
array (nothing allocated) gives the results of tw_batch_run, on two calls in a row (the second
with a smaller batch, so the static array holds the first call's contents when it starts); and a
batch whose launch needs more scratch than the array holds is refused."""
import os

import pytest

import util


def _world(name):
    import twins_amd
    by = {c[0]: c for c in util.cases()}
    _, fl, ref, args = by[name]
    return twins_amd.World([a.replace("topo/", util.GOLDEN + "/topo/") for a in args] + [f"--flows={util.GOLDEN}/{fl}"])


def _results(b, n):
    return [(b.fct(i), b.counters(i)) for i in range(n)]


@pytest.mark.gpu
def test_noalloc_equals_batch_run():
    import twins_amd
    big = [_world(n) for n in ("k8_paper-dctcp-base", "k8_a2a30-newreno-spray", "ls_incast16-newreno-base", "ft8f_a2a30-dctcp-base")]
    small = [_world("k8_incast16-cubic-base")]
    for label, ws, blocks in (("4 worlds", big, 32), ("1 world", small, 8)):
        print(f"noalloc: {label}, {blocks} workgroups", flush=True)
        ref = twins_amd.Batch(ws, blocks=blocks)
        ref.run()
        want = _results(ref, len(ws))
        ref.close()
        b = twins_amd.Batch(ws, blocks=blocks)
        st = b.run_noalloc()
        assert st.worlds_done == len(ws)
        assert _results(b, len(ws)) == want, label
        b.close()
    # a launch beyond the static array: refused, nothing run
    b = twins_amd.Batch(big[:1], blocks=64)
    with pytest.raises(twins_amd.TwError, match="static scratch"):
        b.run_noalloc()
    b.close()
