"""The warm-start network rounds its inputs and activations to bfloat16 with
the hardware conversion (v_cvt_pk_bf16_f32).  The oracle (oracle/mlp.py)
rounds the bit pattern to nearest even; tests/bf16_check.hip compares the two
on the device for all 2^32 float patterns (a NaN only has to stay a NaN)."""
import os
import subprocess

import pytest

from tests import build_checks


@pytest.mark.gpu
def test_hardware_bf16_rounding_matches_oracle_rounding(gpu):
    exe = build_checks.path("bf16_check")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run `python tests/build_checks.py`")
    print("running bf16_check", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and " 0 wrong" in r.stdout
