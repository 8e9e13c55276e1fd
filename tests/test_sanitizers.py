"""The host code under sanitizers (no GPU): the threaded tree builder, its
fast-pass hierarchy build and refit, and the scene-file reader, built with
AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer
(walkray_amd/host/Makefile.sanitize) and run on a heightfield, a degenerate
triangle soup and the Cornell scene files; any report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_host_code_under_asan_ubsan_and_tsan(tmp_path):
    env = dict(os.environ, OUT=str(tmp_path))
    r = subprocess.run(["make", "-s", "-f", "host/Makefile.sanitize", "run", f"OUT={tmp_path}"],
                       cwd=os.path.join(ROOT, "walkray_amd"), capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.count("sanitize_test: all checks passed") == 2
