"""The stencils' correctly rounded square root (sqrt_rn in lps_kernels.hip)
corrects the hardware estimate by a midpoint test.  This pins the test itself:
over every float in [2^-96, 2^64) and every estimate within one ulp of the
root, the corrected value must be the IEEE sqrtf result (C, OpenMP, a few
seconds).  The device code is pinned by the bit-exact GPU parity tests."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_sqrt_midpoint_correction_is_exact(tmp_path):
    exe = str(tmp_path / "sqrt_rn_check")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "sqrt_rn_check.c"), "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert int(r.stdout.split()[0]) > 2_000_000_000
