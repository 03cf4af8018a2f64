"""The oracle and the host-only XYZ reader under AddressSanitizer and
UndefinedBehaviorSanitizer (make -C oracle sanitize, CPU only): every oracle
entry point on five molecules in both Fock modes, the reader on a good, a
malformed, a missing and a concatenated file.  The sanitizers abort on any
finding (-fno-sanitize-recover=all)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_and_reader_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "sanitize: ok" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
