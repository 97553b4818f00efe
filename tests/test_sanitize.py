"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
Netpbm / A.bin readers (gm_netpbm.cpp), the superpixel merge levels
(gm_merge.cpp) and the CPU oracle (oracle/oracle.c), built with gcc/g++ only
(tests/cpp/sanitize) and run on well-formed and malformed inputs."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "cpp", "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no gcc/g++")
def test_host_code_under_asan_ubsan(tmp_path):
    out = tmp_path / "bin"
    b = subprocess.run(["make", "-s", "-C", SAN, f"OUT={out}"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    # verify_asan_link_order=0: the environment may preload a library ahead of the ASan runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(out / "host_sanitize"), str(tmp_path)], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host sanitizer run: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
