"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only):
the oracle's whole path on ragged and degenerate shapes, and the library's
dump reader / SHA-256 against every single-byte change and truncation of a
file (tests/sanitize/)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("gcc") or not shutil.which("g++") or not shutil.which("make"):
        pytest.skip("no host toolchain")
    r = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in r.stderr.lower() or "sanitizer" in r.stderr.lower():
            pytest.skip("sanitizer runtime not available")
        raise AssertionError(r.stdout + r.stderr)
    return os.path.join(HERE, "_build")


def test_oracle_under_sanitizers(built):
    r = subprocess.run([os.path.join(built, "oracle_driver")], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "oracle driver: ok" in r.stdout, r.stdout + r.stderr


def test_dump_reader_under_sanitizers(built, tmp_path):
    r = subprocess.run([os.path.join(built, "dump_driver"), str(tmp_path)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "dump driver: ok" in r.stdout, r.stdout + r.stderr
