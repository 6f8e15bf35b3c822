"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/sanitize_check.c calls every oracle entry point on small inputs
(spectral decompositions of random, repeated, zero and plane-strain tensors;
single points with their tangent across the return-map branches; coloring,
color indexes, a committing residual, element tangents and both block
products on a 2 x 2 patch of 4-node elements).  Built with gcc and
-fsanitize=address,undefined (no recovery), it must print "ok" and exit 0:
no out-of-bounds access, leak or undefined behaviour in the checker.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


def test_oracle_is_clean_under_asan_and_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path / "sanitize_check")
    cmd = [gcc, "-std=c99", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-ffp-contract=off", "-fopenmp", "-o", exe, os.path.join(ORACLE, "sanitize_check.c"),
           os.path.join(ORACLE, "ipf_oracle.c"), "-lm"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if b.returncode != 0 and ("asan" in b.stderr.lower() or "ubsan" in b.stderr.lower()):
        pytest.skip("sanitizer runtime not available: " + b.stderr[-300:])
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="2")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stderr[-2000:]
