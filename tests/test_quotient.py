"""CPU: the reciprocal-and-correction quotient used by the device slab and
sphere tests equals IEEE division on its guarded range (tests/quotient_check.c)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_corrected_quotient_equals_ieee_division(tmp_path):
    exe = tmp_path / "quotient_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-pthread", str(HERE / "quotient_check.c"),
                    "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
