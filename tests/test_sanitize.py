"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/cpp/oracle_sanitize_main.c drives every array-walking oracle function
over edge-case grids (one interior cell, single rows and columns, ragged
sizes, both boundary modes, nucleation embeds wider than a cell) with arrays
allocated at exactly nx * ny doubles.  Host code only (gcc)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_under_sanitizers():
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "oracle_sanitize")
        cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer",
               "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
               os.path.join(ROOT, "oracle", "pf_oracle.c"),
               os.path.join(ROOT, "tests", "cpp", "oracle_sanitize_main.c"), "-lm", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        # verify_asan_link_order=0: the process environment may preload other
        # libraries ahead of the sanitizer runtime
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
        assert "calls clean" in r.stdout
