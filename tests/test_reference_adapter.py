"""The reference-side adapter (include/lps_liphen_adapter.hpp) compiled against
the reference's own headers, syntax only: LiPhEn::HipSolver<SPHSolver> and
HipSolver<PCISPHSolver> are instantiated over LiquidPhysics/include.  The
reference's SPHSolver.cpp needs OpenCL headers this image lacks and its Qt
viewer needs Qt, so nothing is linked or run.

The adapter needs one change in the reference, SPHSolver::m_collisionObjects
made protected.  The first check emulates exactly that change by compiling
the reference's headers with their private sections protected; the second
compiles them as they are and requires that this member is the only thing
the adapter cannot reach.
"""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/LiquidPhysics/include"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers not present")

SOURCE = r"""
#include <algorithm>
#include <cmath>
#include <iostream>
#include <type_traits>
#include <unordered_map>
#include <vector>
#if EMULATE_ACCESS
#define private protected
#endif
#include "SPHSolver.h"
#include "PCISPHSolver.h"
#if EMULATE_ACCESS
#undef private
#endif
#include "lps_liphen_adapter.hpp"
template class LiPhEn::HipSolver<LiPhEn::SPHSolver>;
template class LiPhEn::HipSolver<LiPhEn::PCISPHSolver>;
static_assert(!LiPhEn::HipSPHSolver::kPCISPH && LiPhEn::HipPCISPHSolver::kPCISPH, "solver kinds");
int main() {
    LiPhEn::PhysicSolver* s = nullptr;  // the viewer steps it through the base class
    (void)s;
    return 0;
}
"""


def _compile(emulate: bool):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "adapter_check.cpp")
        with open(src, "w") as f:
            f.write(SOURCE)
        cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
               f"-DEMULATE_ACCESS={1 if emulate else 0}", "-I", REF_INCLUDE, "-I", os.path.join(ROOT, "include"), src]
        return subprocess.run(cmd, capture_output=True, text=True)


def test_adapter_compiles_against_reference_headers():
    r = _compile(True)
    assert r.returncode == 0, r.stderr


def test_adapter_needs_only_the_collider_list_made_protected():
    r = _compile(False)
    assert r.returncode != 0
    errors = [ln for ln in r.stderr.splitlines() if re.search(r"\berror\b", ln)]
    print("\n".join(errors), flush=True)
    assert errors and all("m_collisionObjects" in ln for ln in errors)
