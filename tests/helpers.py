"""Shared helpers for the GPU parity tests: run the same state through the HIP
library and through the oracle."""
from __future__ import annotations

import numpy as np

from liquidsimulation_amd.abi import Colliders, Params
from liquidsimulation_amd.solver import solver_for
from oracle import pyoracle as O


def gpu_run(params: Params, col: Colliders, dt: float, st: "O.State", nsteps: int = 1, lanes=None):
    """The library's run from `st`; lanes (1 or 4) fixes the lanes per
    particle of the stencil kernels, None leaves the automatic choice."""
    s = solver_for(params)
    if lanes is not None:
        s.set_tuning(True, 1, 256, lanes)
    s.set_colliders(col)
    s.set_particles(st.x, st.v, st.vh, st.first)
    stats = s.update(dt, nsteps)
    return s, stats


def ulp_diff(a: np.ndarray, b: np.ndarray) -> int:
    """max distance in float32 units in the last place (sign-magnitude aware)."""
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return int(np.max(np.abs(a - b))) if a.size else 0


def rel_err(a, b, scale=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / max(s, 1e-300))
