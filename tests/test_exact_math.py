"""The device replicas of ns-3's floating-point functions (twins_amd/csrc/exact_math.h), built for
the host, against what ns-3 runs on x86-64: libm's pow and GetSeconds summed in x87 long double,
bit for bit over millions of arguments (Cubic's shapes, general ones, the underflow range)."""
import os
import subprocess

import util


def test_exact_math_host(tmp_path):
    exe = str(tmp_path / "emc")
    src = os.path.join(util.ROOT, "tests", "native", "exact_math_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, src], check=True)
    r = subprocess.run([exe, "1000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "pow: 0 of" in r.stdout and "seconds: 0 of" in r.stdout
