"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU: the
initial-condition generators (astro_amd/csrc/ic.cpp) and the oracle (oracle/oracle.c), driven
by tests/cpp/sanitize_host.cpp.  Any invalid access or undefined behaviour aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs gcc/g++")
def test_host_code_under_asan_ubsan(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-g", "-O1"]
    o1 = tmp_path / "oracle.o"
    subprocess.check_call(["gcc", "-std=c11", "-ffp-contract=off", "-mfma", *san, "-c",
                           os.path.join(ROOT, "oracle", "oracle.c"), "-o", str(o1)])
    exe = tmp_path / "sanitize_host"
    subprocess.check_call(["g++", "-std=c++17", "-ffp-contract=off", *san,
                           os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
                           os.path.join(ROOT, "astro_amd", "csrc", "ic.cpp"), str(o1),
                           "-o", str(exe), "-lm"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert out.stdout.strip().endswith("ok")
