"""GPU tests of the launch machinery: the captured-graph step prefix, the
correction lookahead, the correction kernels' workgroup size, the lanes per
particle of the stencil kernels (one, or a quad that splits the candidates and
adds the terms in the one-lane order) and the autotuner change nothing but
timing -- states, iteration counts and seeds are bit-identical to the eager
path."""
import numpy as np
import pytest

from liquidsimulation_amd import scenes
from liquidsimulation_amd.abi import WARM_LEARNED
from liquidsimulation_amd.solver import PCISPHSolver
from oracle import mlp as M
from tests.helpers import ulp_diff

pytestmark = pytest.mark.gpu


def _scene():
    sc = scenes.small_block(10, 14, 8, dt=0.002, max_density_error_ratio=0.001, stall_iterations=4)
    sc.v[:] = np.array([0.6, -0.3, 0.1], np.float32)
    return sc


def _run(sc, steps, use_graphs, lookahead, warm=None, autotune=False, one_call=False, threads=256, lanes=0):
    s = PCISPHSolver(sc.params.replace(warm_start=WARM_LEARNED) if warm is not None else sc.params)
    if warm is not None:
        s.set_warm_start_weights(*warm)
    s.set_colliders(sc.colliders)
    s.set_particles(sc.x, sc.v)
    s.set_tuning(use_graphs, lookahead, threads, lanes)
    tuned = s.autotune(sc.dt, 3) if autotune else None
    its, used = [], []
    for st in (s.update(sc.dt, steps) if one_call else [s.update(sc.dt, 1)[0] for _ in range(steps)]):
        its.append(st.iterations)
        used.append(st.warm_start_used)
    p = s.get_particles()
    s.close()
    return p, its, used, tuned


@pytest.mark.parametrize("seeded", [False, True])
def test_graph_prefix_and_lookahead_match_eager(gpu, seeded):
    sc = _scene()
    warm = None
    if seeded:
        p = M.he_init(3)
        fs = np.concatenate([np.zeros(M.IN), np.ones(M.IN)]).astype(np.float32)
        warm = (p, fs, 0.02)
    ref, its_ref, used_ref, _ = _run(sc, 24, False, 1, warm)
    print("eager iterations:", its_ref, "seeded:", used_ref, flush=True)
    assert max(its_ref) > 3
    for g, la, t, q in ((True, 1, 256, 0), (True, 2, 128, 1), (False, 3, 64, 1), (False, 1, 128, 4),
                        (True, 1, 64, 1), (False, 1, 256, 1), (True, 2, 256, 4)):
        out, its, used, _ = _run(sc, 24, g, la, warm, threads=t, lanes=q)
        assert its == its_ref and used == used_ref, (g, la, t, q)
        for k in ("x", "v", "vhalf"):
            assert ulp_diff(out[k], ref[k]) == 0, (g, la, t, q, k)


def test_autotune_leaves_the_trajectory_unchanged(gpu):
    sc = _scene()
    ref, its_ref, _, _ = _run(sc, 16, False, 1)
    out, its, _, tuned = _run(sc, 16, False, 1, autotune=True)
    print("autotune chose", tuned, flush=True)
    assert tuned["ms_per_step"] > 0 and tuned["lookahead"] in (1, 2, 3)
    assert its == its_ref
    for k in ("x", "v", "vhalf"):
        assert ulp_diff(out[k], ref[k]) == 0, k


@pytest.mark.parametrize("lookahead", [1, 3])
def test_many_steps_in_one_call_match_step_by_step(gpu, lookahead):
    """Steps issued back to back in one call (no host synchronisation between
    them) leave queued lookahead corrections of one step running while the
    host already polls the next step's corrections; the answers must still be
    the current step's."""
    sc = _scene()
    ref, its_ref, _, _ = _run(sc, 24, False, 1)
    for g in (False, True):
        out, its, _, _ = _run(sc, 24, g, lookahead, one_call=True)
        assert its == its_ref, (g, lookahead, its, its_ref)
        for k in ("x", "v", "vhalf"):
            assert ulp_diff(out[k], ref[k]) == 0, (g, lookahead, k)


@pytest.mark.parametrize("method", ["sph", "pcisph"])
@pytest.mark.parametrize("radius_factor", [3.2, 6.4])
def test_quad_lanes_match_one_lane(gpu, method, radius_factor):
    """The quad stencil kernels (density, forces, SPH pressure, predicted
    density, PCISPH pressure) against the one-lane kernels: every state
    bit-identical over 30 steps, and lps_get_tuning reports the lanes in use
    (a quad below 512 particle blocks when automatic).  Kernel radius 6.4 r_p
    puts about 190 candidates in a stencil row, so the rows run several
    windows of 32 and the replaying passes filter all but the first live."""
    from liquidsimulation_amd.abi import METHOD_PCISPH, METHOD_SPH
    sc = scenes.small_block(14, 14, 14, s=0.016, jitter_amp=2e-4, seed=11,
                            method=METHOD_SPH if method == "sph" else METHOD_PCISPH,
                            kernel_radius_factor=radius_factor)
    sc.v[:] = np.array([0.3, -0.8, 0.2], np.float32)
    outs = []
    for lanes in (1, 4, 0):
        s = PCISPHSolver(sc.params) if method == "pcisph" else None
        if s is None:
            from liquidsimulation_amd.solver import solver_for
            s = solver_for(sc.params)
        s.set_colliders(sc.colliders)
        s.set_particles(sc.x, sc.v)
        s.set_tuning(False, 1, 256, lanes)
        assert s.get_tuning()["lanes_per_particle"] == (4 if lanes == 0 else lanes)
        its = [st.iterations for st in s.update(sc.dt, 30)]
        outs.append((s.get_particles(), its, s.get_fields()))
        s.close()
    (ref, its_ref, f_ref) = outs[0]
    print("iterations:", its_ref, flush=True)
    for out, its, f in outs[1:]:
        assert its == its_ref
        for k in ("x", "v", "vhalf"):
            assert ulp_diff(out[k], ref[k]) == 0, k
        for k in ("rho", "pressure", "force"):
            assert ulp_diff(f[k], f_ref[k]) == 0, k


@pytest.mark.parametrize("dims,auto_lanes", [((73, 56, 32), 4), ((64, 64, 32), 1)])
def test_lane_choice_at_the_block_threshold(gpu, dims, auto_lanes):
    """130,816 particles (511 particle blocks) run a quad per particle by
    default and 131,072 (512 blocks) one lane; over 6 PCISPH steps the
    automatic choice and the other setting give the same state, iteration
    counts and fields bit for bit."""
    from liquidsimulation_amd.abi import METHOD_PCISPH
    sc = scenes.small_block(*dims, s=0.016, jitter_amp=2e-4, seed=13, method=METHOD_PCISPH)
    assert sc.n == dims[0] * dims[1] * dims[2]
    outs = []
    for lanes in (0, 5 - auto_lanes):
        s = PCISPHSolver(sc.params)
        s.set_colliders(sc.colliders)
        s.set_particles(sc.x, sc.v)
        s.set_tuning(False, 1, 256, lanes)
        assert s.get_tuning()["lanes_per_particle"] == (auto_lanes if lanes == 0 else lanes)
        its = [st.iterations for st in s.update(sc.dt, 6)]
        outs.append((s.get_particles(), its, s.get_fields()))
        s.close()
    (ref, its_ref, f_ref), (out, its, f) = outs
    print("iterations:", its_ref, flush=True)
    assert its == its_ref
    for k in ("x", "v", "vhalf"):
        assert ulp_diff(out[k], ref[k]) == 0, k
    for k in ("rho", "pressure", "force"):
        assert ulp_diff(f[k], f_ref[k]) == 0, k
