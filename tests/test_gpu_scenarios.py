"""GPU tests of dynamic scenes and the reference's five scenarios.

Particles emitted mid-run are appended on the device after the current
particles (lps_add_particles), colliders moved between steps are re-uploaded
(lps_set_colliders).  The oracle gets the same events: emitted particles
appended to its state arrays (which, like the library's, hold the previous
step's sorted order), the moved colliders passed to its next step.  Both
sides must stay bit-identical, including when the library grows its device
storage and across the scenarios' emission and collider motion.
"""
import numpy as np
import pytest

from liquidsimulation_amd import scenarios as S
from liquidsimulation_amd import scenes
from liquidsimulation_amd.abi import METHOD_PCISPH, METHOD_SPH
from liquidsimulation_amd.solver import solver_for
from tests.helpers import ulp_diff

pytestmark = pytest.mark.gpu


def _append(oracle, st, x, v):
    n = x.shape[0]
    return oracle.State(np.concatenate([st.x, x]).astype(np.float32), np.concatenate([st.v, v]).astype(np.float32),
                        np.concatenate([st.vh, np.zeros((n, 3), np.float32)]),
                        np.concatenate([st.first, np.ones(n, np.uint8)]))


def _assert_same(p, st):
    assert p["x"].shape == st.x.shape
    for k, ref in (("x", st.x), ("v", st.v), ("vhalf", st.vh)):
        assert ulp_diff(p[k], ref) == 0, k
    np.testing.assert_array_equal(p["first"], st.first)


@pytest.mark.parametrize("method", [METHOD_SPH, METHOD_PCISPH])
def test_added_particles_match_oracle_through_growth(gpu, oracle, method):
    """A block steps 4 times; a second block is added (beyond the storage of
    the first, so the device arrays grow), then a third; 12 steps in all."""
    sc = scenes.small_block(8, 8, 8, method=method, dt=0.002)   # 512 particles, one block of storage
    extra1 = scenes.block_lattice(6, 4, 6, 0.02, origin=(0.02, 0.3, 0.02))
    extra2 = scenes.block_lattice(3, 3, 3, 0.02, origin=(0.05, 0.5, 0.05))
    v1 = np.tile(np.array([[0.0, -0.5, 0.1]], np.float32), (extra1.shape[0], 1))
    s = solver_for(sc.params)
    s.set_colliders(sc.colliders)
    s.set_particles(sc.x)
    st = oracle.State.from_scene(sc.x)
    its_g, its_o = [], []
    for k in range(12):
        if k == 4:
            s.add_particles(extra1, v1)
            st = _append(oracle, st, extra1, v1)
        if k == 8:
            s.add_particles(extra2)
            st = _append(oracle, st, extra2, np.zeros_like(extra2))
        its_g.append(s.update(sc.dt, 1)[0].iterations)
        its_o.append(oracle.step_ocl(sc.params, sc.colliders, sc.dt, st).iterations)
    p = s.get_particles()
    print("iterations", its_g, its_o, flush=True)
    assert its_g == its_o
    _assert_same(p, st)
    # ids count on from the first particles
    assert sorted(p["id"].tolist()) == list(range(st.n))
    s.close()


def _run_both(oracle, name, frames, method, **kw):
    params = S.viewer_params(method)
    r = params.particle_radius
    sg, so = S.SCENARIOS[name](r, **kw), S.SCENARIOS[name](r, **kw)
    s, fr = S.run(sg, frames, params)
    x, v = so.init()
    st = oracle.State.from_scene(x, v)
    its = []
    for k in range(frames):
        em, moved = so.update(S.VIEWER_DT)
        if em is not None and em.x.shape[0]:
            st = _append(oracle, st, em.x, em.v)
        if st.n >= S.MIN_PARTICLES:
            its.append(oracle.step_ocl(params, so.colliders, S.VIEWER_DT, st).iterations)
    p = s.get_particles()
    s.close()
    print(f"{name}: particles {fr[0].particles} -> {fr[-1].particles}, emitted {sum(f.emitted for f in fr)}, "
          f"iterations {[f.iterations for f in fr if f.stepped][-5:]}", flush=True)
    assert [f.iterations for f in fr if f.stepped] == its
    _assert_same(p, st)
    return fr


@pytest.mark.parametrize("name,kw", [
    ("dam_break", {}),
    ("waterfall", {}),
    ("water_drops", {"spawn_rate": 30.0}),   # a drop every 4 frames instead of every 121
    ("sphere_waves", {}),
    ("wave_breaker", {}),
])
def test_scenario_sph_matches_oracle(gpu, oracle, name, kw):
    fr = _run_both(oracle, name, 24, METHOD_SPH, **kw)
    if name in ("waterfall", "water_drops"):
        assert fr[-1].particles > fr[0].particles


@pytest.mark.parametrize("name", ["waterfall", "wave_breaker"])
def test_scenario_pcisph_matches_oracle(gpu, oracle, name):
    _run_both(oracle, name, 16, METHOD_PCISPH)


def test_positions_into_device_buffer_match_readback(gpu):
    """Render interop: positions written into a device buffer (a torch tensor
    standing in for a mapped vertex buffer) equal the host readback, for one
    scene and for an ensemble (scenes back to back); host pointers and short
    buffers are rejected before anything runs on the device."""
    import torch

    from liquidsimulation_amd._lib import lib
    from liquidsimulation_amd.solver import Ensemble

    sc = scenes.small_block(9, 7, 5, method=METHOD_PCISPH, dt=0.002)
    s = solver_for(sc.params)
    s.set_colliders(sc.colliders)
    s.set_particles(sc.x)
    s.update(sc.dt, 3)
    s.add_particles(scenes.block_lattice(3, 3, 3, 0.02, origin=(0.05, 0.4, 0.05)))
    s.update(sc.dt, 2)
    ref = s.get_particles()["x"]
    for k in (3, 4, 7):
        t = torch.full((ref.shape[0], k), -1.0, device="cuda", dtype=torch.float32)
        s.positions_into(t)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(t[:, :3].cpu().numpy(), ref)
        assert bool((t[:, 3:] == -1.0).all())
    # rejected: a host pointer, a buffer too small, a stride below 3
    host = np.zeros((ref.shape[0], 4), np.float32)
    assert lib().lps_copy_positions_device(s._ctx, host.ctypes.data, 4, ref.shape[0], host.nbytes) != 0
    small = torch.zeros((ref.shape[0] - 1, 4), device="cuda", dtype=torch.float32)
    assert lib().lps_copy_positions_device(s._ctx, small.data_ptr(), 4, ref.shape[0], small.numel() * 4) != 0
    assert lib().lps_copy_positions_device(s._ctx, small.data_ptr(), 2, ref.shape[0], 1 << 30) != 0
    s.close()
    a, b = scenes.small_block(6, 6, 6, seed=1), scenes.small_block(5, 9, 4, seed=2)
    e = Ensemble(a.params)
    e.set_scenes([a.x, b.x + np.float32(0.0)])
    e.update(a.dt, 2)
    ref = e.get_particles()["x"]
    t = torch.empty((ref.shape[0], 4), device="cuda", dtype=torch.float32)
    e.positions_into(t)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(t[:, :3].cpu().numpy(), ref)
    e.close()
