"""The workgroup block plan's invariants on the host, under AddressSanitizer
and UndefinedBehaviorSanitizer: tests/plan_invariants.cpp built against
cppimpact_amd/csrc/host.cpp (plain C++, no GPU).  Every LDS corner-force
slot and every partial-force slot the element and boundary-node kernels
read is written exactly once per step, for the right node; every local and
device index is in range."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_block_plan_invariants_under_sanitizers(tmp_path):
    exe = str(tmp_path / "plan_invariants")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(ROOT, "tests", "plan_invariants.cpp"),
           os.path.join(ROOT, "cppimpact_amd", "csrc", "host.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "all plans hold" in p.stdout
    assert p.stdout.count("ok ") == 10
    # the 16-bit corner words fit every plan but the fan's 384-element block
    assert p.stdout.count("corner words fit") == 9
    # both partitions are exercised: the cut Hilbert order on the boxes, the
    # grown blocks on the thin cylinder
    assert " Hilbert blocks" in p.stdout and " grown blocks" in p.stdout
