"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): builds
oracle/_asan/selftest (oracle.c with its self-test, `make -C oracle asan`) and runs it. The
self-test calls every oracle entry point on small, odd and edge sizes and checks identities
the mathematics fixes; a sanitizer report or a failed check fails the test.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_selftest_under_asan_and_ubsan():
    b = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], capture_output=True, text=True,
                       timeout=300)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_asan", "selftest")], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "Sanitizer" not in r.stderr
    assert "all checks passed" in r.stdout
