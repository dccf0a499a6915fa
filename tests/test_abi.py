"""CPU tests of the C ABI boundary: the library loads, exports every function
include/lps.h declares, its ctypes mirrors have the C layouts, and argument
errors are reported without a device."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from liquidsimulation_amd import _lib
from liquidsimulation_amd.abi import (CBox, CGridInfo, CParams, CPlane, CSphere, CStepStats, Params,
                                      LPS_ABI_VERSION, LPS_MLP_PARAMS)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "lps.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("lps_params %zu\n", sizeof(lps_params));
  printf("lps_box %zu\n", sizeof(lps_box));
  printf("lps_sphere %zu\n", sizeof(lps_sphere));
  printf("lps_plane %zu\n", sizeof(lps_plane));
  printf("lps_step_stats %zu\n", sizeof(lps_step_stats));
  printf("lps_grid_info %zu\n", sizeof(lps_grid_info));
  printf("lps_tuning %zu\n", sizeof(lps_tuning));
  printf("mlp_params %d\n", LPS_MLP_PARAMS);
  printf("abi %u\n", LPS_ABI_VERSION);
  printf("noalloc_rows %d\n", LPS_NOALLOC_MAX_ROWS);
  { union { float f; unsigned u; } k; k.f = LPS_WARM_KAPPA; printf("kappa_bits %u\n", k.u); }
  F(lps_params, gravity) F(lps_params, particle_radius) F(lps_params, mass_override)
  F(lps_params, min_iterations) F(lps_params, max_density_error_ratio) F(lps_params, warm_start)
  F(lps_params, stall_iterations)
  F(lps_box, half) F(lps_box, boundary) F(lps_sphere, radius) F(lps_sphere, boundary)
  F(lps_plane, offset) F(lps_step_stats, status) F(lps_step_stats, converged)
  F(lps_grid_info, grid_size) F(lps_grid_info, cell_count)
  return 0;
}
"""


@pytest.fixture(scope="module")
def c_layout():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write(PROBE)
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", exe, src])
        out = subprocess.check_output([exe], text=True)
    return {ln.split()[0]: int(ln.split()[1]) for ln in out.splitlines()}


def test_noalloc_rows_match_header(c_layout):
    from liquidsimulation_amd.abi import LPS_NOALLOC_MAX_ROWS
    assert c_layout["noalloc_rows"] == LPS_NOALLOC_MAX_ROWS


def test_warm_kappa_matches_header(c_layout):
    import numpy as np
    from liquidsimulation_amd.abi import WARM_KAPPA
    assert c_layout["kappa_bits"] == int(np.float32(WARM_KAPPA).view(np.uint32))


def test_struct_layouts_match_header(c_layout):
    assert c_layout["lps_params"] == C.sizeof(CParams)
    assert c_layout["lps_box"] == C.sizeof(CBox)
    assert c_layout["lps_sphere"] == C.sizeof(CSphere)
    assert c_layout["lps_plane"] == C.sizeof(CPlane)
    assert c_layout["lps_step_stats"] == C.sizeof(CStepStats)
    from liquidsimulation_amd.abi import CTuning
    assert c_layout["lps_tuning"] == C.sizeof(CTuning)
    assert c_layout["lps_grid_info"] == C.sizeof(CGridInfo)
    assert c_layout["mlp_params"] == LPS_MLP_PARAMS
    assert c_layout["abi"] == LPS_ABI_VERSION
    for key, off in c_layout.items():
        if "." not in key:
            continue
        tname, field = key.split(".")
        cls = {"lps_params": CParams, "lps_box": CBox, "lps_sphere": CSphere, "lps_plane": CPlane,
               "lps_step_stats": CStepStats, "lps_grid_info": CGridInfo}[tname]
        assert getattr(cls, field).offset == off, key


def test_oracle_params_mirror_public_params():
    from oracle import pyoracle  # noqa: F401  (the oracle takes the same CParams)
    hdr = open(os.path.join(ROOT, "oracle", "orc.h")).read()
    body = hdr[hdr.index("typedef struct orc_params {"):hdr.index("} orc_params;")]
    import re
    body = re.sub(r"/\*.*?\*/", "", body)
    names = [ln.strip().rstrip(";").split()[-1].split("[")[0] for ln in body.splitlines()[1:] if ln.strip()]
    assert names == [f[0] for f in CParams._fields_]


def test_library_exports_every_header_function():
    L = _lib.lib()
    names = _lib.header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # and nothing on the ABI is left undeclared
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.LIB_PATH], text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln and ln.split()[-1].startswith("lps_")}
    assert exported == set(names)


def _has_gpu():
    try:
        from liquidsimulation_amd.solver import device_info
        device_info()
        return True
    except Exception:
        return False


def test_argument_errors_without_device():
    L = _lib.lib()
    vp = C.c_void_p()
    bad = Params().to_c()
    bad.abi_version = 99
    assert L.lps_create(C.byref(bad), C.byref(vp)) == 1  # LPS_ERR_INVALID
    bad = Params(particle_radius=-1.0).to_c()
    assert L.lps_create(C.byref(bad), C.byref(vp)) == 1
    assert L.lps_step(None, C.c_float(0.001), 1, None) == 1
    assert L.lps_destroy(None) == 1
    assert L.lps_status_string(5) == b"LPS_ERR_GRID_OVERFLOW"
    if not _has_gpu():
        good = Params().to_c()
        assert L.lps_create(C.byref(good), C.byref(vp)) == 2  # LPS_ERR_NO_DEVICE: caller falls back


def test_solver_fails_loudly_without_library(monkeypatch):
    """No CPU fallback: a missing library raises."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/liblps.so")
    with pytest.raises(_lib.LibraryMissing):
        _lib.lib()


def test_index_thresholds_reproduce_sqrt_and_division():
    """The stencils find the table index min(trunc(sqrtf(s)/step), 1000) of a
    squared distance s from thresholds instead of a square root and a
    division.  For every index boundary the threshold and its float
    predecessor fall on either side; random squared distances (below h^2 and
    near the boundaries) get the index the IEEE square root and division give."""
    L = _lib.lib()
    for r_p, factor in ((0.00390625, 3.2), (0.0125, 3.2), (0.0025, 3.2), (0.017, 2.69)):
        p = Params(particle_radius=r_p, kernel_radius_factor=factor).to_c()
        thr = np.zeros(1002, np.float32)
        assert L.lps_index_thresholds(C.byref(p), thr) == 0
        h = np.float32(np.float32(factor) * np.float32(r_p))
        step = np.float32(h / np.float32(1000))

        def index(s):
            s = np.asarray(s, np.float32)
            return np.minimum(np.trunc(np.sqrt(s) / step), 1000).astype(np.int64)

        assert thr[0] == 0 and np.isinf(thr[1001])
        k = np.arange(1, 1001)
        finite = np.isfinite(thr[1:1001])
        assert finite[:999].all()  # every index below 1000 is reachable under h^2
        t = thr[1:1001][finite]
        kk = k[finite]
        assert np.all(index(t) >= kk)
        assert np.all(index(np.nextafter(t, np.float32(0))) < kk)
        assert np.all(np.diff(thr[1:1001][finite]) > 0)
        rng = np.random.default_rng(int(r_p * 1e6))
        s = np.concatenate([rng.uniform(0, 1, 200000).astype(np.float32) * h * h,
                            np.nextafter(t, np.float32(0)), t, np.nextafter(t, np.float32(1))]).astype(np.float32)
        via = np.searchsorted(thr[1:1001], s, side="right")  # number of k >= 1 with thr[k] <= s
        np.testing.assert_array_equal(via, index(s))
