"""The C ABI library: it loads, exports every entry point include/twins.h and include/twins_rdma.h
declare, builds worlds on the host, and refuses to simulate without a GPU (it never falls back to
the CPU)."""
import ctypes
import os
import re

import pytest

import util

HDRS = [os.path.join(util.ROOT, "include", h) for h in ("twins.h", "twins_rdma.h")]


def declared():
    names = set()
    for h in HDRS:
        text = open(h).read()
        names |= set(re.findall(r"^\s*(?:int|void|const char \*)\s*(trd_\w+|tw_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    import twins_amd
    L = twins_amd.lib()
    names = declared()
    assert len(names) >= 20 and any(n.startswith("trd_") for n in names)
    for n in names:
        assert hasattr(L, n), n


def test_world_build_on_host():
    import twins_amd
    g = util.GOLDEN
    w = twins_amd.World(["--k=8", f"--flows={g}/flows/k8_paper.csv", "--tcp=dctcp"])
    assert (w.info.hosts, w.info.switches, w.info.ports_per_switch, w.info.flows) == (128, 80, 8, 128)
    assert w.info.links == 128 + 80 * 8
    w2 = twins_amd.World(["--topo=graph", f"--topofile={g}/topo/fattree8.txt", f"--flows={g}/flows/ft8f_a2a30.csv",
                          "--fail=300000:s32:s0,500000:s64:s32,800000:s32:s0:up"])
    assert w2.info.hosts == 128 and w2.info.switches == 80


@pytest.mark.parametrize("args,msg", [(["--k=6", f"--flows={util.GOLDEN}/flows/k8_paper.csv"], "power of two"), (["--k=8"], "--flows"),
                                      (["--k=8", "--flows=/nonexistent.csv"], "cannot open"), (["--bogus=1"], "unknown argument")])
def test_world_errors(args, msg):
    import twins_amd
    with pytest.raises(twins_amd.TwError, match=re.escape(msg)):
        twins_amd.World(args)


def test_no_cpu_fallback():
    """Without a GPU the batch refuses to exist (the model runs on the GPU only)."""
    import twins_amd
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    w = twins_amd.World(["--k=8", f"--flows={util.GOLDEN}/flows/k8_paper.csv"])
    with pytest.raises(twins_amd.TwError, match="no GPU"):
        twins_amd.Batch([w])
