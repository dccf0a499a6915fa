"""ctypes mirrors of the C ABI structs declared in include/lps.h.

The layouts are the public contract of liblps.so; the field order follows
lps_params / lps_box / lps_sphere / lps_plane / lps_step_stats / lps_grid_info.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, asdict
from typing import List, Sequence

LPS_ABI_VERSION = 1
LPS_TABLE_SIZE = 1001
LPS_MLP_IN = 16
LPS_MLP_HIDDEN = 64
LPS_MLP_PARAMS = (LPS_MLP_IN * LPS_MLP_HIDDEN + LPS_MLP_HIDDEN + LPS_MLP_HIDDEN * LPS_MLP_HIDDEN
                  + LPS_MLP_HIDDEN + LPS_MLP_HIDDEN + 1)

METHOD_SPH = 0
METHOD_PCISPH = 1
WARM_NONE = 0
WARM_LEARNED = 1
WARM_LEARNED_ALWAYS = 2
# rows accepted by lps_mlp_grad_noalloc
LPS_NOALLOC_MAX_ROWS = 65536
# seed envelope factor of the learned warm start (LPS_WARM_KAPPA, a float)
WARM_KAPPA = 0.85
# relative energy tolerance of the warm start's fallback (LPS_WARM_ENERGY_TOL)
WARM_ENERGY_TOL = 1e-4

STATUS_NAMES = {
    0: "LPS_OK",
    1: "LPS_ERR_INVALID",
    2: "LPS_ERR_NO_DEVICE",
    3: "LPS_ERR_OOM",
    4: "LPS_ERR_NONFINITE",
    5: "LPS_ERR_GRID_OVERFLOW",
    6: "LPS_ERR_HIP",
    7: "LPS_ERR_STATE",
}


class CParams(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("method", C.c_int32),
        ("gravity", C.c_float * 3),
        ("particle_radius", C.c_float),
        ("kernel_radius_factor", C.c_float),
        ("rest_density", C.c_float),
        ("stiffness", C.c_float),
        ("negative_pressure_factor", C.c_float),
        ("viscosity", C.c_float),
        ("surface_tension", C.c_float),
        ("surface_tension_threshold", C.c_float),
        ("restitution", C.c_float),
        ("friction", C.c_float),
        ("mass_override", C.c_float),
        ("min_iterations", C.c_int32),
        ("max_iterations", C.c_int32),
        ("max_density_error_ratio", C.c_float),
        ("warm_start", C.c_int32),
        ("stall_iterations", C.c_int32),
    ]


class CBox(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("half", C.c_float * 3), ("boundary", C.c_int32)]


class CSphere(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("boundary", C.c_int32)]


class CPlane(C.Structure):
    _fields_ = [("normal", C.c_float * 3), ("offset", C.c_float)]


class CStepStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("max_density_error_ratio", C.c_float),
        ("warm_start_used", C.c_int32),
        ("status", C.c_int32),
        ("converged", C.c_int32),
    ]


class CTrainCfg(C.Structure):
    """lps_mlp_train_cfg."""
    _fields_ = [
        ("epochs", C.c_uint32),
        ("batch", C.c_uint32),
        ("lr", C.c_float),
        ("lr_final", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
        ("reinit", C.c_int32),
        ("refit_scale", C.c_int32),
        ("over_weight", C.c_float),
        ("seed", C.c_uint64),
    ]


class CTuning(C.Structure):
    """lps_tuning."""
    _fields_ = [
        ("use_graphs", C.c_int32),
        ("lookahead", C.c_uint32),
        ("ms_per_step", C.c_float),
        ("correction_threads", C.c_uint32),
        ("lanes_per_particle", C.c_uint32),
    ]


class CGridInfo(C.Structure):
    _fields_ = [
        ("bbox_min", C.c_float * 3),
        ("bbox_max", C.c_float * 3),
        ("grid_offset", C.c_float * 3),
        ("grid_spacing", C.c_float),
        ("grid_size", C.c_uint32 * 3),
        ("cell_count", C.c_uint32),
    ]


@dataclass
class Params:
    """Solver parameters with the reference constructor defaults
    (LiquidPhysics/src/SPHSolver.cpp:11-23, PCISPHSolver.cpp:6-7)."""

    method: int = METHOD_PCISPH
    gravity: Sequence[float] = (0.0, -9.81, 0.0)
    particle_radius: float = 0.017
    kernel_radius_factor: float = 2.688
    rest_density: float = 998.29
    stiffness: float = 3.0
    negative_pressure_factor: float = 0.0
    viscosity: float = 3.5
    surface_tension: float = 0.0728
    surface_tension_threshold: float = 7.065
    restitution: float = 0.5
    friction: float = 1.0
    mass_override: float = 0.0
    min_iterations: int = 4
    max_iterations: int = 4
    max_density_error_ratio: float = 0.1
    warm_start: int = WARM_NONE
    stall_iterations: int = 0

    def to_c(self) -> CParams:
        c = CParams()
        c.abi_version = LPS_ABI_VERSION
        c.method = int(self.method)
        for k in range(3):
            c.gravity[k] = float(self.gravity[k])
        for name in ("particle_radius", "kernel_radius_factor", "rest_density", "stiffness",
                     "negative_pressure_factor", "viscosity", "surface_tension",
                     "surface_tension_threshold", "restitution", "friction", "mass_override",
                     "max_density_error_ratio"):
            setattr(c, name, float(getattr(self, name)))
        c.min_iterations = int(self.min_iterations)
        c.max_iterations = int(self.max_iterations)
        c.warm_start = int(self.warm_start)
        c.stall_iterations = int(self.stall_iterations)
        return c

    def replace(self, **kw) -> "Params":
        d = asdict(self)
        d.update(kw)
        return Params(**d)


@dataclass
class Box:
    center: Sequence[float]
    half: Sequence[float]
    boundary: bool = False


@dataclass
class Sphere:
    center: Sequence[float]
    radius: float
    boundary: bool = False


@dataclass
class Plane:
    normal: Sequence[float]
    offset: float


@dataclass
class Colliders:
    boxes: List[Box] = field(default_factory=list)
    spheres: List[Sphere] = field(default_factory=list)
    planes: List[Plane] = field(default_factory=list)

    def c_arrays(self):
        nb, ns, npl = len(self.boxes), len(self.spheres), len(self.planes)
        boxes = (CBox * max(nb, 1))()
        spheres = (CSphere * max(ns, 1))()
        planes = (CPlane * max(npl, 1))()
        for i, b in enumerate(self.boxes):
            for k in range(3):
                boxes[i].center[k] = float(b.center[k])
                boxes[i].half[k] = float(b.half[k])
            boxes[i].boundary = 1 if b.boundary else 0
        for i, s in enumerate(self.spheres):
            for k in range(3):
                spheres[i].center[k] = float(s.center[k])
            spheres[i].radius = float(s.radius)
            spheres[i].boundary = 1 if s.boundary else 0
        for i, p in enumerate(self.planes):
            for k in range(3):
                planes[i].normal[k] = float(p.normal[k])
            planes[i].offset = float(p.offset)
        return boxes, nb, spheres, ns, planes, npl
