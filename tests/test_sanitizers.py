"""The host-side planner, operand-reuse planner, tile grouping and CPU plan executor built with
AddressSanitizer + UndefinedBehaviorSanitizer (g++, no HIP) and run over plans of every shape
class the executor handles (tests/cpp/sanitize_host.cpp): no sanitizer report, exact results."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tiled-mm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_planner_and_cpu_executor_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", CSRC,
           os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
           os.path.join(CSRC, "tmm_plan.cpp"), os.path.join(CSRC, "tmm_cpu_exec.cpp"), "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "SANITIZE HOST OK" in r.stdout, (r.stdout[-3000:], r.stderr[-6000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
