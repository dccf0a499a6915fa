"""The pressure-force stencils divide three numerators by one distance d with
one shared reciprocal and Markstein's correction (liquidsimulation_amd/csrc/lps_div.h)
where that is exact, and with IEEE division elsewhere.  This pins the method:

* CPU: for every d in [2^-78, 2^40) and every reciprocal estimate within one
  ulp of 1/d, the Newton step gives the correctly rounded 1/d, except for an
  all-ones significand with the lower estimate; random quotients match IEEE
  division (C, OpenMP, a few seconds).
* GPU: with the hardware estimate (never that lower one), the reciprocal of
  every such d equals 1.0f/d, and the corrected quotient equals the IEEE
  quotient for slices of significand pairs.  The exhaustive run over all 2^46 pairs is
  `tests/_bin/div_rn_check pairs 0 8388608` (about a minute on one MI355X).
"""
import os
import shutil
import subprocess

import pytest

from tests import build_checks

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_reciprocal_step_is_exact_cpu(tmp_path):
    exe = str(tmp_path / "div_rn_check")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "div_rn_check.c"), "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert int(r.stdout.split()[0]) > 1_900_000_000


def _run(*args):
    exe = build_checks.path("div_rn_check")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run `python tests/build_checks.py`")
    print("running", os.path.basename(exe), *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    return r


@pytest.mark.gpu
def test_device_reciprocal_matches_ieee(gpu):
    r = _run("rcp")
    assert r.returncode == 0 and " 0 wrong" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("a0,a1", [(0, 2048), (4194304 - 1024, 4194304 + 1024), (8388608 - 2048, 8388608)])
def test_device_quotients_match_ieee(gpu, a0, a1):
    r = _run("pairs", str(a0), str(a1))
    assert r.returncode == 0 and " 0 wrong" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_table_index_quotient_is_exact_cpu(tmp_path, oracle):
    """The pressure passes take the lookup index trunc(d / div) as the
    corrected quotient with div's correctly rounded reciprocal (table_index_q).
    For the table step of every scene the tests and the benchmark run, every
    float distance below 1001 steps gives the IEEE quotient's index."""
    from liquidsimulation_amd import scenarios, scenes
    params = [scenes.small_block().params, scenes.sph_dam_break_16k().params,
              scenes.pcisph_dam_break_524k().params, scenes.pcisph_obstacles_2m().params,
              scenarios.viewer_params()]
    divs = sorted({oracle.derived(p)[1] for p in params})
    exe = str(tmp_path / "div_rn_check")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "div_rn_check.c"), "-lm"], check=True)
    r = subprocess.run([exe, *[float(d).hex() for d in divs]], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(" 0 wrong") == len(divs)
