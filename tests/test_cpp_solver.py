"""The C++ host layer (include/lps_solver.hpp): the reference's SPHSolver /
PCISPHSolver interface over the C ABI.  tests/cpp_solver_check.cpp steps a
small dam break through it (setters, addParticle, colliders, update, readback,
particles added mid-run) and compares every step with the oracle, SPH and
PCISPH, bit for bit."""
import os
import shutil
import subprocess

import pytest

from tests import build_checks

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "lps_solver.hpp"\n'
                   'int main() { lps::PCISPHSolver s; s.setMinIterations(3); s.setGravity({0.f, -9.81f, 0.f});\n'
                   '  s.addParticle({0.1f, 0.1f, 0.1f}); return s.getParticleCount() == 1 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)


@pytest.mark.gpu
def test_cpp_layer_steps_match_oracle(gpu):
    exe = build_checks.path("cpp_solver_check")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run `python tests/build_checks.py`")
    print("running cpp_solver_check", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok")
