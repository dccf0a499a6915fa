"""Device energy reduction (lps_get_energy) and the warm start's energy check.

The library sums every particle's kinetic (m |v|^2 / 2) and potential
(-m g.x, per-scene gravity) energy per block in the integration (and in the
particle upload), and per scene in double; the warm start's fallback compares
consecutive steps' energies so that a seed which adds energy is switched off.
No counterpart in the reference.
"""
import numpy as np
import pytest

from liquidsimulation_amd import scenes
from liquidsimulation_amd.abi import (LPS_MLP_IN, LPS_MLP_PARAMS, METHOD_PCISPH, METHOD_SPH, WARM_LEARNED,
                                      WARM_LEARNED_ALWAYS)
from liquidsimulation_amd.solver import Ensemble, PCISPHSolver, solver_for

pytestmark = pytest.mark.gpu


def _host_energy(mass, gravity, x, v):
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    return 0.5 * mass * float(np.sum(v * v)), -mass * float(np.sum(x @ np.asarray(gravity, np.float64)))


@pytest.mark.parametrize("method", [METHOD_SPH, METHOD_PCISPH])
def test_energy_matches_readback(gpu, method):
    """After the upload and after steps: per-block FP32 sums reduced in double
    agree with a float64 sum over the read-back state to 1e-6."""
    sc = scenes.small_block(12, 16, 10, method=method, dt=0.002)
    s = solver_for(sc.params)
    s.set_colliders(sc.colliders)
    x0 = sc.x
    v0 = np.random.default_rng(1).normal(0, 0.3, x0.shape).astype(np.float32)
    s.set_particles(x0, v0)
    mass = s.getParticleMass()
    for steps in (0, 1, 7):
        if steps:
            s.update(sc.dt, steps)
        p = s.get_particles()
        ke, pe = s.get_energy()
        hk, hp = _host_energy(mass, sc.params.gravity, p["x"], p["v"])
        print(f"after {steps} steps: ke {ke[0]:.6g} vs {hk:.6g}, pe {pe[0]:.6g} vs {hp:.6g}", flush=True)
        assert abs(ke[0] - hk) <= 1e-6 * max(hk, 1e-12) + 1e-12
        assert abs(pe[0] - hp) <= 1e-6 * abs(hp)
    s.close()


def test_energy_per_scene_of_an_ensemble(gpu):
    """Three scenes with their own gravity: each scene's energies equal the
    host sums over its own particles."""
    members = [scenes.small_block(8, 10, 8, seed=s, dt=0.002) for s in (1, 2, 3)]
    gravity = [(0.0, -9.81, 0.0), (1.5, -9.7, 0.0), (0.0, -9.6, -2.0)]
    e = Ensemble(members[0].params)
    e.set_scenes([m.x for m in members], gravity=gravity)
    for k, m in enumerate(members):
        e.set_scene_colliders(k, m.colliders)
    e.update(0.002, 3)
    ke, pe = e.get_energy()
    mass = e.getParticleMass()
    parts = e.get_particles()
    for k, sl in enumerate(e.scene_slices()):
        hk, hp = _host_energy(mass, gravity[k], parts["x"][sl], parts["v"][sl])
        print(f"scene {k}: ke {ke[k]:.6g} vs {hk:.6g}, pe {pe[k]:.6g} vs {hp:.6g}", flush=True)
        assert abs(ke[k] - hk) <= 1e-6 * hk
        assert abs(pe[k] - hp) <= 1e-6 * abs(hp)
    e.close()


def _envelope_net():
    """A network whose output is a constant far above any pressure: the seed
    is then the envelope's bound max(p1, kappa * p_prev) everywhere."""
    p = np.zeros(LPS_MLP_PARAMS, np.float32)
    p[-1] = 1.0e4
    fs = np.concatenate([np.zeros(LPS_MLP_IN), np.ones(LPS_MLP_IN)]).astype(np.float32)
    return p, fs


def test_fallback_switches_off_a_seed_that_adds_energy(gpu):
    """The 16k PCISPH dam break with the bound alone as the seed: at kappa =
    0.85 the pressure of particles held by the walls feeds on itself from step
    to step.  Forced on (no fallback) the scene gains energy; with the
    fallback the energy check switches the seed off (it keeps fewer
    corrections, so the iteration rule alone would keep it) and the total
    energy stays within 1 % of its start."""
    sc = scenes.pcisph_dam_break_16k()
    p, fs = _envelope_net()
    res = {}
    for mode in (WARM_LEARNED_ALWAYS, WARM_LEARNED):
        s = PCISPHSolver(sc.params.replace(warm_start=mode))
        s.set_colliders(sc.colliders)
        s.set_particles(sc.x, sc.v)
        s.set_warm_start_weights(p, fs, 1.0)
        ke, pe = s.get_energy()
        e0 = float(ke[0] + pe[0])
        emax, used = e0, 0
        for _ in range(20):
            st = s.update(sc.dt, 20)
            used += sum(x.warm_start_used for x in st)
            ke, pe = s.get_energy()
            emax = max(emax, float(ke[0] + pe[0]))
        res[mode] = (emax / e0, used)
        s.close()
    print("max energy / start, seeded steps (of 400): forced", res[WARM_LEARNED_ALWAYS], "fallback",
          res[WARM_LEARNED], flush=True)
    assert res[WARM_LEARNED_ALWAYS][0] > 1.05
    assert res[WARM_LEARNED][0] <= 1.01
    assert res[WARM_LEARNED][1] < 200
