"""Helpers of the parity tests: the golden cases, running the oracle and the GPU command line,
comparing outputs column by column as the reference's validation/check_exact.sh does."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE = os.path.join(ROOT, "oracle", "tcp_oracle")
TWINS = os.path.join(ROOT, "twins_amd", "bin", "twins")


def oracle_bin():
    """The oracle binary, built here with make if it is missing (a CPU build, no GPU needed)."""
    if not os.path.exists(ORACLE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    return ORACLE


def cases():
    """(name, flows, ns-3 output, argument list) of every stored golden case."""
    out = []
    for line in open(os.path.join(GOLDEN, "manifest.txt")):
        if line.startswith("#") or not line.strip():
            continue
        w, t, v, fl, ref, args = line.rstrip("\n").split("|")
        out.append((f"{w}-{t}-{v}", fl, ref, args.split() + [f"--tcp={t}"]))
    return out


def run(binary, args, out_csv, extra=(), timeout=600):
    """Runs a `twins`-style binary from tests/golden (relative paths); returns its stdout+stderr.
    A GPU binary that fails in any way ends the test session: nothing more runs on the GPU."""
    gpu = binary == TWINS
    r = subprocess.run([binary] + list(args) + [f"--out={out_csv}"] + list(extra), cwd=GOLDEN, capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode < 0 or r.returncode in (124, 134, 137, 139) or (gpu and r.returncode != 0):
        import pytest
        # a crash, abort or kill: nothing more runs on the GPU in this session
        pytest.exit(f"{os.path.basename(binary)} {' '.join(args)} ended with {r.returncode}: {r.stderr[-2000:]}", returncode=3)
    if r.returncode != 0:
        raise RuntimeError(f"{os.path.basename(binary)} {' '.join(args)} failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return r.stdout + r.stderr


def read_cols(path, cols):
    rows = []
    with open(path) as f:
        next(f)
        for line in f:
            v = line.rstrip("\n").split(",")
            rows.append(tuple(v[c] for c in cols))
    return rows


def same_times(ref_csv, out_csv):
    """(flows identical in established_ns and completed_ns, flows) -- check_exact.sh's measure."""
    a, b = read_cols(ref_csv, (4, 5)), read_cols(out_csv, (4, 5))
    assert len(a) == len(b), f"{len(a)} reference rows, {len(b)} output rows"
    return sum(x == y for x, y in zip(a, b)), len(a)


def undecided(log):
    return "warning" in log and "undecided" in log


def read_qtrace(path):
    """Records of a queue trace (the reference's --qtrace format): (ts, node, dev, src, dst, kind,
    seq-or-ack, qlen) per line."""
    out = []
    with open(path) as f:
        for line in f:
            v = line.split()
            out.append((int(v[0]), int(v[1]), int(v[2]), int(v[3]), int(v[4]), v[5], int(v[6]), int(v[7])))
    return out


def packet_events(log):
    return int(log.split("packet-events ")[1].split()[0])
