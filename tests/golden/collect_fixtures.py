"""Copies the reference's stored test data used by the parity tests into tests/golden/.

Fixtures (data files only): the flow files and topology files the reference's validation runs
read (data/flows, data/topo), and the outputs of sequential ns-3 those runs are compared with
(data/ns3/ref: src,dst,flow_size,start_ns,established_ns,completed_ns,hash_data,hash_ack[,retx]).
The selection covers every feature of the TCP model the reference validates
(validation/check_all.sh): the three TCPs, FatTree / leaf-spine / topology files, start-jittered
flows, sizes that are not MSS multiples, persistent connections with and without FIN,
per-packet ECMP, and link failures.

Run once in a checkout that has the reference at /root/reference:
    python3 tests/golden/collect_fixtures.py
"""
import os
import shutil
import sys

REF = "/root/reference/data"
HERE = os.path.dirname(os.path.abspath(__file__))

# (workload, tcps, variant); variant base|fin|spray|sprayfin|jit*  (validation/check_exact.sh)
T3 = ("dctcp", "newreno", "cubic")
CASES = [
    ("k8_paper", T3, "base"), ("k8_fullsync", T3, "base"), ("k8_fullexp", ("dctcp",), "base"),
    ("k8_a2a30", T3, "base"), ("k8_incast16", T3, "base"), ("k8_a2a50s", ("dctcp",), "base"),
    ("k8_paper", T3, "jit"), ("k8_a2a30", ("newreno",), "jit2"),
    ("k8_a2a30b", T3, "base"), ("k8_incastb", ("dctcp",), "base"),
    ("k8_r30", T3, "base"), ("k8_rpc", ("dctcp",), "base"), ("k8_r40p1", ("cubic",), "base"),
    ("k8_r30", T3, "fin"), ("k8_a2a30", ("dctcp",), "fin"), ("k8_incast16", ("cubic",), "fin"),
    ("k8_paper", T3, "spray"), ("k8_a2a30", T3, "spray"), ("k8_r30", T3, "sprayfin"),
    ("ls_paper", T3, "base"), ("ls_a2a30", ("dctcp", "cubic"), "base"), ("ls_incast16", ("newreno",), "base"),
    ("ls_a2a30", ("dctcp",), "spray"), ("ls_r30", ("dctcp",), "base"),
    ("jf32_a2a30", T3, "base"), ("jf32_a2a30", T3, "spray"), ("jf32_incast", ("dctcp",), "base"),
    ("torus6_a2a30", ("dctcp", "newreno"), "base"), ("df4x9_a2a30", ("cubic",), "spray"),
    ("ft8f_a2a30", T3, "base"), ("ft8f_a2a30", T3, "spray"), ("jf32f_a2a50", ("dctcp",), "base"),
    ("torus6f_a2a30", ("newreno",), "base"),
    ("k16_paper", ("dctcp",), "base"), ("k16_a2a30", ("dctcp", "cubic"), "base"),
    ("k16_incast16", ("newreno",), "base"), ("k16_a2a50s", T3, "base"),
]
TCP = {"dctcp": "TcpDctcp", "newreno": "TcpNewReno", "cubic": "TcpCubic"}


def ref_name(w, t, v):
    if v == "base":
        return f"{w}_{TCP[t]}.csv"
    if v in ("fin", "spray", "sprayfin"):
        return f"{w}_{TCP[t]}_{v}.csv"
    return f"{w}_{TCP[t]}_{v}.csv"


def flow_name(w, v):
    return f"{w}.csv" if v in ("base", "fin", "spray", "sprayfin") else f"{w}_{v}.csv"


def main():
    os.makedirs(os.path.join(HERE, "flows"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "topo"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "ns3"), exist_ok=True)
    manifest = []
    for w, tcps, v in CASES:
        fl = flow_name(w, v)
        shutil.copy(os.path.join(REF, "flows", fl), os.path.join(HERE, "flows", fl))
        topo = os.path.join(REF, "flows", w + ".csv.topo")
        args = f"--k={w.split('_')[0][1:]}" if w.startswith("k") else ""
        if os.path.exists(topo):
            line = open(topo).read().split("# twins ", 1)[1].strip()
            for tok in line.split():
                if tok.startswith("--topofile=data/topo/"):
                    tf = tok.split("/")[-1]
                    shutil.copy(os.path.join(REF, "topo", tf), os.path.join(HERE, "topo", tf))
            args = line.replace("--topofile=data/topo/", "--topofile=topo/")
        for t in tcps:
            rn = ref_name(w, t, v)
            shutil.copy(os.path.join(REF, "ns3", "ref", rn), os.path.join(HERE, "ns3", rn))
            extra = {"base": "", "fin": " --fin=1", "spray": " --spray=1", "sprayfin": " --spray=1 --fin=1"}.get(v, "")
            manifest.append(f"{w}|{t}|{v}|flows/{fl}|ns3/{rn}|{args}{extra}")
    with open(os.path.join(HERE, "manifest.txt"), "w") as f:
        f.write("# workload|tcp|variant|flows|ns-3 output|twins arguments (paths relative to tests/golden)\n")
        f.write("\n".join(manifest) + "\n")
    print(len(manifest), "cases")


if __name__ == "__main__":
    sys.exit(main())
