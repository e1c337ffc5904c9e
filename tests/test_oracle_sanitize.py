"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(`make -C oracle sanitize`, host code only) and run over a small scene by
tests/orc_sanitize_main.c: every entry point, every hit group, texture
format and filter, ray flag and thread count, with the BVH queries checked
against brute force. A sanitizer finding aborts the run."""
import os
import subprocess

from conftest import REPO


def test_oracle_under_sanitizers():
    oracle = os.path.join(REPO, "oracle")
    b = subprocess.run(["make", "-C", oracle, "sanitize"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([os.path.join(oracle, "_build", "orc_sanitize")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "oracle sanitize run: ok" in r.stdout
