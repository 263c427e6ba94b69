"""The oracle (oracle/mgp_oracle.c, mgp_oracle4.c) built with
AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile `asan`) and
run over every entry point, every side type, both homogeneous modes and the
three BC modes (oracle/asan_driver.c): no out-of-bounds access, use after
free, leak or undefined behaviour.  Host code only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def test_oracle_under_address_and_ub_sanitizers():
    try:
        subprocess.run(["make", "-C", ORACLE, "asan"], check=True, capture_output=True, text=True, timeout=300)
    except subprocess.CalledProcessError as e:
        if "asan" in (e.stderr or "") and "cannot find" in (e.stderr or ""):
            pytest.skip("the C compiler has no sanitizer runtime")
        raise
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_asan", "oracle_asan")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle sanitizer run: checksum" in r.stdout
    assert "runtime error" not in r.stderr
