"""The McICA kernel converts MT19937 draws to [0, 1] as the reference's
random_number_t does (y / 4294967295 in FP64), but without a division.  The
replacement is exact for every 32-bit draw; this checks all 2^32 of them on
the CPU (a few seconds with OpenMP)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_division_free_uniform_is_exact_for_all_draws():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "chk")
        # -ffp-contract=off: only the explicit fma() calls may fuse
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-o", exe,
                        os.path.join(HERE, "mt_uniform_exhaustive.c"), "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600)
    assert out.stdout.strip() == "0"
