"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
`make -C oracle sanitize` builds oracle/san_check.c (which compiles
tcr_oracle.c into itself) with -fsanitize=address,undefined and no recovery,
and the binary drives every oracle entry point over edge-case shapes (1-3
dimensions, all modes, colliding hash-map lengths, points on and beyond the
domain bounds, empty sets, search range 2).  Any sanitizer report aborts it;
it also checks the hashed counts against the O(N^2) brute force."""
import os
import re
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_clean_under_asan_and_ubsan():
    if shutil.which("make") is None or shutil.which("cc") is None:
        pytest.skip("no make / C compiler")
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    if b.returncode != 0 and ("asan" in b.stderr or "ubsan" in b.stderr):
        pytest.skip("the C compiler has no sanitizer runtime: " + b.stderr[-300:])
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "san_check")], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    m = re.search(r"(\d+) cases \((\d+) against the brute force\), (\d+) failures", r.stdout)
    assert m and int(m.group(3)) == 0
    assert int(m.group(1)) >= 250 and int(m.group(2)) >= 150
