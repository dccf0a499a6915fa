"""The reference viewer's five scenarios, headless, on the HIP solver.

LiquidSimulation (the Qt viewer) steps one scenario at a time: each frame it
calls the scenario's updateScenario(dt), which may emit particles or move a
collider, then steps the solver once the world holds at least 512 particles
(LiquidSimulation::stepSimulation, LiquidSimulation/src/LiquidSimulation.cpp:175-181).
This module restates the scenarios' initial particles and colliders
(initScenario) and their per-frame updates (updateScenario) in FP32 with the
reference's expression orders, at the viewer's default slider values
(LiquidSimulation.cpp:9-77, restoreDefaultParameters), and runs them:

* emitted particles are appended to the device state (`lps_add_particles`),
  after the particles of the last step in their sorted order, as the reference
  appends new particle objects to its solver's list;
* moved colliders are re-uploaded (`lps_set_colliders`), with no particle
  upload and no graph re-capture.

The only deviation: Water Drops places its drops with qrand() seeded by the
wall clock (LiquidSimulation.cpp:6-7); here the draws come from a seeded
splitmix64 sequence, so runs are reproducible.

    python -m liquidsimulation_amd.scenarios dam_break --steps 500 [--pcisph]
"""
from __future__ import annotations

import argparse
import json
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from .abi import METHOD_PCISPH, METHOD_SPH, Box, Colliders, Params, Sphere
from .scenes import spawn_cube, splitmix64

f32 = np.float32


def slider(step: float, value: int) -> float:
    """A viewer parameter: slider step x slider value, a float product."""
    return float(f32(step) * f32(value))


def viewer_params(method: int = METHOD_SPH) -> Params:
    """The viewer's defaults (LiquidSimulation.cpp:9-77): time step 83 x 1e-4 s,
    gravity (0, -98 x 0.1, 0), particle radius 17 x 1e-3, kernel radius factor
    269 x 0.01, rest density 998, viscosity 35 x 0.1, stiffness 30 x 0.1,
    negative pressure factor 0, surface tension 73 x 1e-3 with threshold
    71 x 0.1, restitution 50 x 0.01, friction 100 x 0.01.  PCISPH keeps the
    reference solver's loop knobs (PCISPHSolver.cpp:6-7): at most 4
    corrections, stopping below a 10 % density error."""
    return Params(
        method=method,
        gravity=(slider(0.1, 0), slider(0.1, -98), slider(0.1, 0)),
        particle_radius=slider(0.001, 17),
        kernel_radius_factor=slider(0.01, 269),
        rest_density=slider(1.0, 998),
        stiffness=slider(0.1, 30),
        negative_pressure_factor=slider(0.01, 0),
        viscosity=slider(0.1, 35),
        surface_tension=slider(0.001, 73),
        surface_tension_threshold=slider(0.1, 71),
        restitution=slider(0.01, 50),
        friction=slider(0.01, 100),
        min_iterations=1,
        max_iterations=4,
        max_density_error_ratio=0.1,
    )


VIEWER_DT = slider(0.0001, 83)


# ------------------------------------------------------------- emitters --
def _lattice_offset(radius: float, r: f32) -> Tuple[f32, int, f32]:
    step = f32(1.6) * r
    size = ((f32(2.0) * f32(radius)) - (f32(2.0) * r)) / step
    res = int(size) + 1
    rem = size - f32(int(size))
    return step, res, rem


def spawn_sphere(position, sphere_radius: float, particle_radius: float) -> np.ndarray:
    """SPHParticleEmitter::spawnSphere (SPHParticleEmitter.cpp:45-75): lattice
    points of step 1.6 r whose distance from the centre plus r is at most the
    sphere radius, x-outer / z-inner order."""
    r = f32(particle_radius)
    R = f32(sphere_radius)
    step, res, rem = _lattice_offset(sphere_radius, r)
    off = (f32(1.0) + rem / f32(2.0)) * r
    pos = np.asarray(position, np.float32)
    out = []
    for x in range(res):
        for y in range(res):
            for z in range(res):
                lat = np.array([f32(x) * step, f32(y) * step, f32(z) * step], np.float32)
                d = (np.float32(off) + lat) - R
                dist = f32(math.sqrt(float(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2])))) + r
                if dist <= R:
                    out.append(((pos - R) + off) + lat)
    return np.asarray(out, np.float32).reshape(-1, 3)


def spawn_circle_z(position, circle_radius: float, particle_radius: float) -> np.ndarray:
    """SPHParticleEmitter::spawnCircleZ (SPHParticleEmitter.cpp:137-162): a disc
    in the xy plane.  The reference's position offset term is written
    Vector3D((a, b, 0.f) * r), a comma expression whose value is 0 (the
    :150 bug), so the lattice starts at position - (R, R, 0) exactly; the
    distance test uses the intended (1 + rem/2) r offset."""
    r = f32(particle_radius)
    R = f32(circle_radius)
    step, res, rem = _lattice_offset(circle_radius, r)
    off = (f32(1.0) + rem / f32(2.0)) * r
    pos = np.asarray(position, np.float32)
    corner = pos - np.array([R, R, f32(0.0)], np.float32)
    out = []
    for x in range(res):
        for y in range(res):
            dx = f32(f32(off + f32(x) * step) - R)
            dy = f32(f32(off + f32(y) * step) - R)
            dist = f32(math.sqrt(float(f32(f32(dx * dx) + f32(dy * dy)) + f32(0.0)))) + r
            if dist <= R:
                lat = np.array([f32(x) * step, f32(y) * step, f32(0.0)], np.float32)
                out.append((corner + np.zeros(3, np.float32)) + lat)
    return np.asarray(out, np.float32).reshape(-1, 3)


# ------------------------------------------------------------ scenarios --
@dataclass
class Emission:
    x: np.ndarray
    v: np.ndarray


@dataclass
class Scenario:
    """initScenario + updateScenario of one reference scenario."""
    name: str
    r: float
    time: f32 = f32(0.0)
    colliders: Colliders = field(default_factory=Colliders)

    def init(self) -> Tuple[np.ndarray, np.ndarray]:
        raise NotImplementedError

    def update(self, dt: float) -> Tuple[Optional[Emission], bool]:
        """Returns the particles emitted this frame (or None) and whether the
        colliders changed."""
        return None, False


def _sin_wave(t: f32, frequency: f32, amplitude: f32) -> f32:
    # sin(t * f * 2.f * M_PI) * a: float products, then double (M_PI), back to float
    arg = float(f32(f32(t * frequency) * f32(2.0))) * math.pi
    return f32(math.sin(arg) * float(amplitude))


class WaterDrops(Scenario):
    """WaterDropsScenario.cpp: a shallow pool in a tall boundary box; every
    1/spawn_rate seconds a drop (spawnSphere of radius drop_size) falls from
    y = 0.5 at a random (x, z)."""

    def __init__(self, r: float, seed: int = 7, spawn_rate: float = slider(0.1, 10),
                 drop_size: float = slider(0.01, 10)):
        super().__init__("water_drops", r)
        self.spawn_rate = f32(spawn_rate)
        self.drop_size = f32(drop_size)
        self._draws = (splitmix64(seed, 4096) % np.uint64(101)).astype(np.int64)
        self._k = 0

    def init(self):
        x = spawn_cube((0.0, -0.7, 0.0), (0.6, 0.1, 0.6), self.r)
        self.colliders = Colliders(boxes=[Box(center=(0.0, 0.0, 0.0), half=(0.6, 0.8, 0.6), boundary=True)])
        self.time = f32(0.0)
        return x, np.zeros_like(x)

    def _rand(self) -> f32:
        k = self._draws[self._k % self._draws.size]
        self._k += 1
        return f32(f32(f32(f32(k) * f32(0.01)) * f32(1.2)) - f32(0.6))

    def update(self, dt):
        self.time = f32(self.time + f32(dt))
        if self.time >= f32(f32(1.0) / self.spawn_rate):
            xr, zr = self._rand(), self._rand()
            xr = f32(xr + self.drop_size) if xr < f32(0.0) else f32(xr - self.drop_size)
            zr = f32(zr + self.drop_size) if zr < f32(0.0) else f32(zr - self.drop_size)
            x = spawn_sphere((xr, 0.5, zr), float(self.drop_size), self.r)
            self.time = f32(0.0)
            return Emission(x, np.zeros_like(x)), False
        return None, False


class Waterfall(Scenario):
    """WaterfallScenario.cpp: a pool behind two obstacle steps; a disc of
    particles (spawnCircleZ of radius inflow_size) enters through the back wall
    at inflow_speed along +z every 1.6 r / inflow_speed seconds."""

    def __init__(self, r: float, inflow_speed: float = slider(0.1, 18), inflow_size: float = slider(0.01, 10)):
        super().__init__("waterfall", r)
        self.speed = f32(inflow_speed)
        self.size = f32(inflow_size)

    def init(self):
        x = spawn_cube((-0.75, -0.2, 0.0), (0.45, 0.1, 0.6), self.r)
        self.colliders = Colliders(boxes=[
            Box(center=(0.0, 0.0, 0.0), half=(1.2, 0.6, 0.6), boundary=True),
            Box(center=(-0.2, -0.35, 0.0), half=(0.1, 0.25, 0.6), boundary=False),
            Box(center=(-0.7, -0.45, 0.0), half=(0.5, 0.15, 0.6), boundary=False),
        ])
        self.time = f32(0.0)
        return x, np.zeros_like(x)

    def update(self, dt):
        r = f32(self.r)
        self.time = f32(self.time + f32(dt))
        if self.time >= f32(f32(r * f32(1.6)) / self.speed):
            x = spawn_circle_z((-0.8, 0.3, float(f32(f32(-0.6) + r))), float(self.size), self.r)
            v = np.zeros_like(x)
            v[:, 2] = self.speed
            self.time = f32(0.0)
            return Emission(x, v), False
        return None, False


class SphereWaves(Scenario):
    """SphereWavesScenario.cpp: a cube of fluid inside a boundary sphere whose
    radius oscillates as 1 + a sin(2 pi f t), around an obstacle column."""

    def __init__(self, r: float, frequency: float = slider(0.1, 6), amplitude: float = slider(0.01, 10)):
        super().__init__("sphere_waves", r)
        self.freq = f32(frequency)
        self.amp = f32(amplitude)

    def init(self):
        x = spawn_cube((0.0, -0.25, 0.0), (0.45, 0.45, 0.45), self.r)
        self.colliders = Colliders(
            boxes=[Box(center=(0.0, -0.65, 0.0), half=(0.1, 0.5, 0.1), boundary=False)],
            spheres=[Sphere(center=(0.0, 0.0, 0.0), radius=1.0, boundary=True)])
        self.time = f32(0.0)
        return x, np.zeros_like(x)

    def update(self, dt):
        s = _sin_wave(self.time, self.freq, self.amp)
        self.time = f32(self.time + f32(dt))
        sp = self.colliders.spheres[0]
        self.colliders.spheres[0] = Sphere(center=sp.center, radius=float(f32(f32(1.0) + s)), boundary=True)
        return None, True


class DamBreak(Scenario):
    """DamBreakScenario.cpp: a column of fluid released towards an obstacle box
    and a half-buried obstacle sphere; no per-frame update."""

    def __init__(self, r: float):
        super().__init__("dam_break", r)

    def init(self):
        x = spawn_cube((-0.6, -0.2, 0.0), (0.4, 0.4, 0.6), self.r)
        self.colliders = Colliders(
            boxes=[Box(center=(0.0, 0.0, 0.0), half=(1.0, 0.6, 0.6), boundary=True),
                   Box(center=(0.4, -0.2, -0.4), half=(0.2, 0.6, 0.2), boundary=False)],
            spheres=[Sphere(center=(0.4, -0.25, 0.6), radius=0.3, boundary=False)])
        self.time = f32(0.0)
        return x, np.zeros_like(x)


class WaveBreaker(Scenario):
    """WaveBreakerScenario.cpp: a ball of fluid in a box whose -x wall moves as
    a piston (half width 1.2 - s/2, centre s/2, s = a sin(2 pi f t)) against
    three obstacle columns."""

    def __init__(self, r: float, frequency: float = slider(0.1, 10), amplitude: float = slider(0.01, 10)):
        super().__init__("wave_breaker", r)
        self.freq = f32(frequency)
        self.amp = f32(amplitude)

    def init(self):
        x = spawn_sphere((0.6, 0.0, 0.0), 0.6, self.r)
        self.colliders = Colliders(boxes=[
            Box(center=(0.0, 0.0, 0.0), half=(1.2, 0.6, 0.6), boundary=True),
            Box(center=(-0.2, -0.2, 0.0), half=(0.1, 0.4, 0.1), boundary=False),
            Box(center=(-0.2, -0.2, -0.4), half=(0.1, 0.4, 0.1), boundary=False),
            Box(center=(-0.2, -0.2, 0.4), half=(0.1, 0.4, 0.1), boundary=False),
        ])
        self.time = f32(0.0)
        return x, np.zeros_like(x)

    def update(self, dt):
        s = _sin_wave(self.time, self.freq, self.amp)
        self.time = f32(self.time + f32(dt))
        b = self.colliders.boxes[0]
        half = (float(f32(f32(1.2) - f32(s / f32(2.0)))), b.half[1], b.half[2])
        center = (float(f32(s / f32(2.0))), b.center[1], b.center[2])
        self.colliders.boxes[0] = Box(center=center, half=half, boundary=True)
        return None, True


SCENARIOS: Dict[str, Callable[[float], Scenario]] = {
    "water_drops": WaterDrops,
    "waterfall": Waterfall,
    "sphere_waves": SphereWaves,
    "dam_break": DamBreak,
    "wave_breaker": WaveBreaker,
}

MIN_PARTICLES = 512  # the viewer steps the solver only from 512 particles on


@dataclass
class Frame:
    particles: int
    stepped: bool
    emitted: int
    iterations: int


def run(scenario: Scenario, steps: int, params: Params, dt: float = VIEWER_DT, solver=None,
        on_frame: Optional[Callable[[int, Scenario, Optional[Emission], bool], None]] = None) -> Tuple[object, List[Frame]]:
    """LiquidSimulation::stepSimulation for `steps` frames on the HIP solver
    (a fresh one unless given).  on_frame(k, scenario, emission, colliders_changed)
    is called after each frame's update and before its solver step."""
    from .solver import solver_for
    s = solver or solver_for(params)
    x, v = scenario.init()
    s.set_colliders(scenario.colliders)
    s.set_particles(x, v)
    n = x.shape[0]
    frames: List[Frame] = []
    for k in range(steps):
        em, moved = scenario.update(dt)
        if em is not None and em.x.shape[0]:
            s.add_particles(em.x, em.v)
            n += em.x.shape[0]
        if moved:
            s.set_colliders(scenario.colliders)
        if on_frame is not None:
            on_frame(k, scenario, em, moved)
        its = 0
        stepped = n >= MIN_PARTICLES
        if stepped:
            its = s.update(dt, 1)[0].iterations
        frames.append(Frame(n, stepped, 0 if em is None else em.x.shape[0], its))
    return s, frames


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("scenario", choices=sorted(SCENARIOS))
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--pcisph", action="store_true", help="PCISPH instead of the viewer's default SPH")
    args = ap.parse_args()
    params = viewer_params(METHOD_PCISPH if args.pcisph else METHOD_SPH)
    sc = SCENARIOS[args.scenario](params.particle_radius)
    t0 = time.perf_counter()
    s, frames = run(sc, args.steps, params)
    p = s.get_particles()
    wall = time.perf_counter() - t0
    s.close()
    ps = sum(f.particles for f in frames if f.stepped)
    print(json.dumps({"scenario": args.scenario, "method": "PCISPH" if args.pcisph else "SPH", "frames": args.steps,
                      "dt": VIEWER_DT, "particles_start": frames[0].particles, "particles_end": frames[-1].particles,
                      "emitted": sum(f.emitted for f in frames), "wall_s": wall,
                      "particle_steps_per_s_wall": ps / wall if wall > 0 else None,
                      "finite": bool(np.isfinite(p["x"]).all())}))


if __name__ == "__main__":
    main()
