"""Seeded dam-break scenes for the benchmark workloads.

The reference builds its scenes in the Qt scenarios with
SPHParticleEmitter::spawnCube (LiquidPhysics/src/Particles/SPHParticleEmitter.cpp:13-43,
used by LiquidSimulation/src/Scenarios/DamBreakScenario.cpp:9-35).  The workloads
here place a block of nx*ny*nz particles on the lattice x = s*(i+1) (FP32,
i = 0..nx-1, likewise y and z) in the min corner of a BOUNDARY box, in the
emitter's loop order (x outer, y, z inner), and add i.i.d. uniform jitter.

The lattice step is s = 1.6 * r_p, the step of the reference's mass lattice
(SPHSolver.cpp:485), so an interior particle of the unjittered block has the
rest density by construction of the particle mass.

Jitter generator (documented so any consumer can regenerate it): splitmix64
with state seed, one output per component in (particle, axis) order;
u = (z >> 40) * 2^-24 in FP32 (exact), jitter = amp * (2u - 1) in FP32.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .abi import Box, Colliders, Params, Plane, Sphere, METHOD_PCISPH, METHOD_SPH

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count outputs of splitmix64 starting from state `seed` (uint64 wraparound)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    z = splitmix64(seed, count)
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def jitter(seed: int, count: int, amp: float) -> np.ndarray:
    u = uniform01(seed, count)
    return np.float32(amp) * (np.float32(2.0) * u - np.float32(1.0))


def spawn_cube(position, half_sizes, particle_radius) -> np.ndarray:
    """SPHParticleEmitter::spawnCube (SPHParticleEmitter.cpp:13-43) in FP32: a
    lattice of step 1.6 r_p filling the cube, centred by half the remainder,
    in x-outer / z-inner order."""
    f = np.float32
    r = f(particle_radius)
    step = f(1.6) * r
    pos = np.asarray(position, np.float32)
    half = np.asarray(half_sizes, np.float32)
    size = ((f(2.0) * half) - (f(2.0) * r)) / step
    res = size.astype(np.int32) + 1
    rem = size - size.astype(np.int32).astype(np.float32)
    off = (f(1.0) + rem / f(2.0)) * r
    base = (pos - half) + off
    ix = np.arange(res[0]).astype(np.float32) * step
    iy = np.arange(res[1]).astype(np.float32) * step
    iz = np.arange(res[2]).astype(np.float32) * step
    X, Y, Z = np.meshgrid(ix, iy, iz, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    return (base + lat).astype(np.float32)


def block_lattice(nx: int, ny: int, nz: int, s: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """[n,3] FP32 lattice origin + s*(i+1), x-outer / z-inner order."""
    s32 = np.float32(s)
    ix = (np.arange(nx, dtype=np.float32) + np.float32(1.0)) * s32 + np.float32(origin[0])
    iy = (np.arange(ny, dtype=np.float32) + np.float32(1.0)) * s32 + np.float32(origin[1])
    iz = (np.arange(nz, dtype=np.float32) + np.float32(1.0)) * s32 + np.float32(origin[2])
    X, Y, Z = np.meshgrid(ix, iy, iz, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)


@dataclass
class Scene:
    name: str
    params: Params
    colliders: Colliders
    x: np.ndarray               # [n,3] float32
    v: np.ndarray               # [n,3] float32
    dt: float
    steps: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def with_params(self, **kw) -> "Scene":
        """The same scene with some fluid/solver parameters replaced."""
        return Scene(name=self.name, params=self.params.replace(**kw), colliders=self.colliders, x=self.x,
                     v=self.v, dt=self.dt, steps=self.steps, meta=dict(self.meta))


def _fluid_params(s: float, method: int, **kw) -> Params:
    # r_p = s / 1.6 so the reference's mass lattice step 1.6 r_p equals s;
    # h = 2 s = 3.2 r_p.  Fluid constants: rest density 1000, stiffness 3.0,
    # viscosity 0.01; surface tension 0.0728 / threshold 7.065, restitution 0.5,
    # friction 1.0 are the reference constructor defaults (SPHSolver.cpp:11-21).
    p = Params(
        method=method,
        gravity=(0.0, -9.81, 0.0),
        particle_radius=float(np.float32(s) / np.float32(1.6)),
        kernel_radius_factor=3.2,
        rest_density=1000.0,
        stiffness=3.0,
        negative_pressure_factor=0.0,
        viscosity=0.01,
        surface_tension=0.0728,
        surface_tension_threshold=7.065,
        restitution=0.5,
        friction=1.0,
        min_iterations=3,
        max_iterations=50,
        max_density_error_ratio=0.01,
        stall_iterations=3,
    )
    return p.replace(**kw) if kw else p


def dam_break(nx: int, ny: int, nz: int, s: float, box=(1.6, 1.0, 0.8), jitter_amp: float = 0.0,
              seed: int = 0, method: int = METHOD_PCISPH, dt: float = 0.001, steps: int = 100,
              name: Optional[str] = None, **param_kw) -> Scene:
    x = block_lattice(nx, ny, nz, s)
    if jitter_amp > 0.0:
        x = (x + jitter(seed, x.size, jitter_amp).reshape(x.shape)).astype(np.float32)
    bx, by, bz = (float(b) for b in box)
    col = Colliders(boxes=[Box(center=(bx / 2, by / 2, bz / 2), half=(bx / 2, by / 2, bz / 2), boundary=True)])
    p = _fluid_params(s, method, **param_kw)
    return Scene(name=name or f"dam_break_{nx}x{ny}x{nz}", params=p, colliders=col, x=x,
                 v=np.zeros_like(x), dt=dt, steps=steps,
                 meta={"nx": nx, "ny": ny, "nz": nz, "spacing": s, "box": box, "jitter": jitter_amp,
                       "seed": seed})


def sph_dam_break_16k(seed: int = 1) -> Scene:
    """Plain SPH dam break, 16,000 particles: 20x40x20 at 0.02 m in a 1.6x1.0x0.8 m box,
    dt 0.002 s, 500 steps, +-1e-4 m jitter."""
    return dam_break(20, 40, 20, 0.02, jitter_amp=1e-4, seed=seed, method=METHOD_SPH, dt=0.002,
                     steps=500, name="sph_dam_break_16k")


def pcisph_dam_break_16k(seed: int = 1) -> Scene:
    """The 16,000-particle dam break (20x40x20 at 0.02 m, dt 0.002 s) run with
    PCISPH at the 524k scene's tolerances (1 %, 3..50 corrections)."""
    return dam_break(20, 40, 20, 0.02, jitter_amp=1e-4, seed=seed, method=METHOD_PCISPH, dt=0.002,
                     steps=500, name="pcisph_dam_break_16k")


def pcisph_dam_break_524k(seed: int = 2) -> Scene:
    """PCISPH dam break, 524,288 particles: 64x128x64 at 0.00625 m in the same box,
    dt 0.001 s, 2000 steps, +-1e-5 m jitter, tolerance 1 %, 3..50 corrections."""
    return dam_break(64, 128, 64, 0.00625, jitter_amp=1e-5, seed=seed, method=METHOD_PCISPH,
                     dt=0.001, steps=2000, name="pcisph_dam_break_524k")


def obstacle_spheres(seed: int, count: int, block_max: float, box=(2.0, 1.0, 1.0)) -> List[Sphere]:
    """`count` non-overlapping OBSTACLE spheres, radius U[0.05, 0.15], centres
    drawn by rejection from splitmix64(seed) in x in [block_max + r + 0.02, 1.3],
    y in [r, 0.45], z in [r, box_z - r]."""
    out: List[Sphere] = []
    u = uniform01(seed, 4 * 1000)
    k = 0
    while len(out) < count and k + 4 <= u.size:
        r = 0.05 + 0.10 * float(u[k])
        lo = block_max + r + 0.02
        cx = lo + (1.3 - lo) * float(u[k + 1])
        cy = r + (0.45 - r) * float(u[k + 2])
        cz = r + (box[2] - 2 * r) * float(u[k + 3])
        k += 4
        ok = all(math.dist((cx, cy, cz), sp.center) > r + sp.radius + 0.02 for sp in out)
        if ok:
            out.append(Sphere(center=(cx, cy, cz), radius=r, boundary=False))
    return out


def ramp_plane(angle_deg: float = 20.0, x0: float = 1.3) -> Plane:
    """A plane tilted `angle_deg` from the floor, rising towards +x from (x0, 0, z):
    normal (-sin a, cos a, 0), fluid side n.x >= offset."""
    a = math.radians(angle_deg)
    n = (-math.sin(a), math.cos(a), 0.0)
    return Plane(normal=n, offset=n[0] * x0)


def pcisph_obstacles_2m(seed: int = 3) -> Scene:
    """PCISPH with static obstacles, 2,097,152 particles: 128^3 at 0.004 m in a
    2.0x1.0x1.0 m box, 3 obstacle spheres and a 20 degree plane, dt 0.0005 s, 1000 steps."""
    sc = dam_break(128, 128, 128, 0.004, box=(2.0, 1.0, 1.0), jitter_amp=1e-5, seed=seed,
                   method=METHOD_PCISPH, dt=0.0005, steps=1000, name="pcisph_obstacles_2m")
    block_max = 0.004 * 129
    sc.colliders.spheres = obstacle_spheres(seed + 1000, 3, block_max, box=(2.0, 1.0, 1.0))
    sc.colliders.planes = [ramp_plane(20.0, 1.3)]
    return sc


def small_block(nx=8, ny=8, nz=8, s=0.02, jitter_amp=1e-4, seed=5, method=METHOD_PCISPH, dt=0.002,
                box=None, **kw) -> Scene:
    """A small dam break for unit and parity tests (box twice the block's extent)."""
    if box is None:
        box = (4 * s * (nx + 1), 2 * s * (ny + 1), 2 * s * (nz + 1))
    return dam_break(nx, ny, nz, s, box=box, jitter_amp=jitter_amp, seed=seed, method=method, dt=dt,
                     steps=10, name=f"small_{nx}x{ny}x{nz}", **kw)


@dataclass
class EnsembleScene:
    """Independent scenes stepped together (shared fluid parameters)."""
    name: str
    params: Params
    members: List[Scene]            # per-scene positions, velocities, colliders
    gravity: List[tuple]            # per-scene gravity
    dt: float
    steps: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(sum(m.n for m in self.members))


def _log_uniform_counts(u, n_min, n_max, total):
    """Target counts n_min * (n_max / n_min)^(c u_k): log-uniform in
    [n_min, n_min (n_max / n_min)^c], with c <= 1 found by bisection so that
    the targets sum to `total` (c = 1 when the full range sums to less)."""
    r = math.log(n_max / n_min)
    f = lambda c: sum(n_min * math.exp(r * c * x) for x in u)
    if f(1.0) <= total:
        return 1.0
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if f(mid) < total else (lo, mid)
    return 0.5 * (lo + hi)


def ensemble_64(seed: int = 4, count: int = 64, n_min: int = 8192, n_max: int = 131072,
                total: int = 2_000_000) -> EnsembleScene:
    """Batched ensemble: `count` dam breaks with particle counts log-uniform
    between n_min and at most n_max, about `total` in all, drawn from
    splitmix64(seed).  A log-uniform draw over the whole of [8k, 131k] has a
    mean of 44k per scene (2.8M for 64), so the range's upper end is narrowed
    by a common exponent c (see _log_uniform_counts) to make the total about
    2M.  Per-scene block dims (ny = 2 nx, nz in [nx/2, 2 nx]), gravity tilted
    by up to +-10 degrees about z and x, one obstacle sphere in the collapse
    path; spacing 0.00625 m and the 524k scene's fluid and solver parameters,
    dt 0.001 s, 500 steps."""
    s_ = 0.00625
    u = uniform01(seed, 8 * count).astype(np.float64).reshape(count, 8)
    c = _log_uniform_counts(u[:, 0], n_min, n_max, total)
    members, gravity = [], []
    for k in range(count):
        n_t = math.exp(math.log(n_min) + (math.log(n_max) - math.log(n_min)) * c * u[k, 0])
        rz = 0.5 + 1.5 * u[k, 1]                       # nz / nx
        nx = max(4, int(round((n_t / (2.0 * rz)) ** (1.0 / 3.0))))
        ny, nz = 2 * nx, max(4, int(round(rz * nx)))
        bx, by, bz = 4.0 * s_ * (nx + 1), 1.5 * s_ * (ny + 1), s_ * (nz + 3)
        sc = dam_break(nx, ny, nz, s_, box=(bx, by, bz), jitter_amp=1e-5, seed=seed * 1000 + k,
                       method=METHOD_PCISPH, dt=0.001, steps=500, name=f"member_{k}")
        r = (0.15 + 0.15 * u[k, 2]) * min(by, bz)
        sc.colliders.spheres = [Sphere(center=(bx * (0.55 + 0.25 * u[k, 3]), r * (0.6 + 0.8 * u[k, 4]), bz * 0.5),
                                       radius=r, boundary=False)]
        tz, tx = math.radians(20.0 * (u[k, 5] - 0.5)), math.radians(20.0 * (u[k, 6] - 0.5))
        g = (9.81 * math.sin(tz), -9.81 * math.cos(tz) * math.cos(tx), 9.81 * math.cos(tz) * math.sin(tx))
        members.append(sc)
        gravity.append(tuple(float(v) for v in g))
    return EnsembleScene(name="ensemble_64", params=members[0].params, members=members, gravity=gravity,
                         dt=0.001, steps=500, meta={"seed": seed, "count": count, "range_exponent": c})


WORKLOADS = {
    "sph_dam_break_16k": sph_dam_break_16k,
    "pcisph_dam_break_16k": pcisph_dam_break_16k,
    "pcisph_dam_break_524k": pcisph_dam_break_524k,
    "pcisph_obstacles_2m": pcisph_obstacles_2m,
    "ensemble_64": ensemble_64,
}
