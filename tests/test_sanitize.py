"""AddressSanitizer + UndefinedBehaviorSanitizer build of the host symbolic
analysis (spdepy_amd/csrc/symbolic.cpp, plain g++, no device code) run over
lattices of every boundary kind, leaf size and the wide-time pattern, with the
invariants the device launches rely on checked on the host."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_symbolic_analysis_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_symbolic")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", os.path.join(HERE, "native", "sanitize_symbolic.cpp"),
           os.path.join(ROOT, "spdepy_amd", "csrc", "symbolic.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert r.stdout.count(" ok") == 7
