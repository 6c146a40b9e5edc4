"""`twins --topo=graph --dump_routes=F` (src/main.cu:466-486): every switch's equal-cost output
interfaces toward every host, in the ns-3 scenario's DUMP_ROUTES format. Checked against what the
topology file fixes: at a host's own switch, the host's port; elsewhere exactly the ports whose
neighbour is one hop closer (breadth-first distance) to the host's switch; ports numbered from 1 in
the order the file's links attach them. Host code only (no GPU: --hash_check stops the run first)."""
import collections
import os
import subprocess

import util


def expected(topo):
    hosts = switches = 0
    ports = collections.defaultdict(list)  # switch -> [neighbour: ("h", i) or ("s", j)]
    for line in open(topo):
        v = line.split()
        if not v:
            continue
        if v[0] == "hosts":
            hosts = int(v[1])
        elif v[0] == "switches":
            switches = int(v[1])
        elif v[0] == "link":
            a, b = (v[1][0], int(v[1][1:])), (v[2][0], int(v[2][1:]))
            if a[0] == "s":
                ports[a[1]].append(b)
            if b[0] == "s":
                ports[b[1]].append(a)
    home = {}
    for x in range(switches):
        for p, n in enumerate(ports[x]):
            if n[0] == "h":
                home[n[1]] = (x, p)
    out = []
    for x in range(switches):
        for h in range(hosts):
            e, hp = home[h]
            if e == x:
                out.append(f"route {hosts + x} {h} {hp + 1}")
                continue
            dist = {e: 0}
            q = collections.deque([e])
            while q:
                y = q.popleft()
                for n in ports[y]:
                    if n[0] == "s" and n[1] not in dist:
                        dist[n[1]] = dist[y] + 1
                        q.append(n[1])
            near = [p + 1 for p, n in enumerate(ports[x]) if n[0] == "s" and dist.get(n[1], -1) == dist[x] - 1]
            out.append(f"route {hosts + x} {h} " + " ".join(map(str, near)))
    return out


def test_dump_routes_fattree8(tmp_path):
    name, fl, ref, args = next(c for c in util.cases() if c[0] == "ft8f_a2a30-dctcp-base")
    topo = [a for a in args if a.startswith("--topofile=")][0].split("=", 1)[1]
    out = tmp_path / "routes.txt"
    r = subprocess.run([util.TWINS, "--topo=graph", f"--topofile={topo}", f"--flows={fl}", f"--dump_routes={out}",
                        f"--ecmp_from={ref}", "--hash_check=1"], cwd=util.GOLDEN, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = out.read_text().splitlines()
    assert got == expected(os.path.join(util.GOLDEN, topo))
    assert len(got) == 128 * 80


def test_dump_routes_ignored_without_topology_file(tmp_path):
    """A FatTree built from --k has no topology file: the reference writes the dump in its graph
    branch only and runs on, and so does this command line (no file, exit 0 at the hash check)."""
    name, fl, ref, args = next(c for c in util.cases() if c[0] == "k8_paper-dctcp-base")
    out = tmp_path / "routes.txt"
    r = subprocess.run([util.TWINS] + args + [f"--flows={fl}", f"--dump_routes={out}", f"--ecmp_from={ref}", "--hash_check=1"],
                       cwd=util.GOLDEN, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert not out.exists()


def test_empty_worlds_file(tmp_path):
    w = tmp_path / "worlds.txt"
    w.write_text("# nothing but a comment\n\n")
    r = subprocess.run([util.TWINS, "--k=4", f"--worlds={w}", f"--dump_routes={tmp_path / 'r.txt'}"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "no worlds" in r.stderr
