"""The reference's dam-break checks (unit_test/unit_test_dam_break.py), run on
the GPU path through the driver, plus multi-substep parity with the oracle and
HIP-graph capture of a substep."""
import math
from dataclasses import replace

import numpy as np
import pytest
import torch

from fluid_simulation_amd.examples.dam_break.scene import DEFAULT_SPEC, ParticleScene, make_scene
from fluid_simulation_amd.examples.dam_break.simulation import SOLVERS, DamBreakSimulation
from tests.cases import oracle_config, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def test_finite_check_detects_invalid_state():
    """unit_test_dam_break.py:39-46"""
    sim = DamBreakSimulation(make_scene(512), "pbf", device="cuda:0")
    sim.check_finite()
    positions = sim.state_0.particle_q.numpy()
    positions[0, 0] = np.nan
    sim.state_0.particle_q.assign(positions)
    with pytest.raises(RuntimeError, match="Non-finite"):
        sim.check_finite()


def test_solvers_share_initial_particles_and_mass():
    """unit_test_dam_break.py:48-57"""
    scene = make_scene(512)
    for method in SOLVERS:
        sim = DamBreakSimulation(scene, method, device="cuda:0")
        np.testing.assert_array_equal(sim.state_0.particle_q.numpy(), scene.positions)
        np.testing.assert_array_equal(sim.state_0.particle_qd.numpy(), np.zeros_like(scene.positions))
        np.testing.assert_allclose(sim.model.particle_radius.numpy(), scene.radius)
        np.testing.assert_allclose(sim.model.particle_mass.numpy(), scene.mass)
        np.testing.assert_allclose(sim.model.gravity.numpy()[0], scene.spec.gravity)


@pytest.mark.parametrize("method", list(SOLVERS))
def test_every_method_advances_complete_frames(method):
    """unit_test_dam_break.py:59-75"""
    scene = make_scene(512)
    sim = DamBreakSimulation(scene, method, device="cuda:0", substeps=4)
    for frame in range(1, 3):
        sim.step()
        assert sim.sim_time == frame / 60.0
        assert sim.last_substeps >= 4
        assert sim.metrics()["finite"]
        x = sim.state_0.particle_q.numpy()
        assert np.all(x >= np.asarray(scene.spec.bound_lo) - 1e-6)
        assert np.all(x <= np.asarray(scene.spec.bound_hi) + 1e-6)
    if method == "wcsph":
        assert sim.total_substeps > 8


def test_underdense_block_falls_without_pressure_attraction():
    """unit_test_dam_break.py:79-91"""
    sim = DamBreakSimulation(make_scene(512), "wcsph", device="cuda:0")
    sim.step()
    assert np.all(sim.state_0.sph.rho.numpy() < 1000.0)
    np.testing.assert_array_equal(sim.state_0.sph.pressure.numpy(), 0.0)
    expected = np.tile(np.asarray(sim.scene.spec.gravity) * sim.frame_dt, (sim.scene.count, 1))
    np.testing.assert_allclose(sim.state_0.particle_qd.numpy(), expected, atol=1e-7)
    signed = DamBreakSimulation(make_scene(512), "wcsph", device="cuda:0", wcsph_clamp_negative_pressure=False)
    signed.step()
    assert np.any(signed.state_0.sph.pressure.numpy() < 0.0)
    assert signed.metrics()["speed_max"] > sim.metrics()["speed_max"]


def test_compressed_particles_repel_and_conserve_momentum():
    """unit_test_dam_break.py:93-107"""
    d = 0.01
    center = np.asarray([0.0, 0.0, 0.5], dtype=np.float32)
    offsets = np.asarray([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float32) * (0.05 * d)
    scene = ParticleScene(replace(DEFAULT_SPEC, gravity=(0.0, 0.0, 0.0), target_count=4), center + offsets,
                          1000.0 * d ** 3, d * 0.5, d ** 3, d)
    sim = DamBreakSimulation(scene, "wcsph", device="cuda:0", wcsph_viscosity=0.0)
    sim._substep(1e-6)
    assert np.all(sim.state_0.sph.rho.numpy() > 1000.0)
    assert np.all(sim.state_0.sph.pressure.numpy() > 0.0)
    velocity = sim.state_0.particle_qd.numpy()
    assert np.all(np.sum(velocity * offsets, axis=1) > 0.0)
    np.testing.assert_allclose(velocity.sum(axis=0), 0.0, atol=1e-6)


def test_cfl_reduction_preserves_maxima_and_rejects_nonfinite_values():
    """unit_test_dam_break.py:109-130"""
    sim = DamBreakSimulation(make_scene(512), "wcsph", device="cuda:0")
    rng = np.random.default_rng(42)
    velocity = rng.normal(size=(sim.scene.count, 3)).astype(np.float32)
    acceleration = (rng.normal(size=velocity.shape) * 20).astype(np.float32)
    sim.state_0.particle_qd.assign(velocity)
    sim.solver._particle_accel.assign(acceleration)
    sim.total_substeps = 1
    speed, accel = sim._cfl_maxima()
    np.testing.assert_allclose([speed, accel], [np.linalg.norm(velocity, axis=1).max(),
                                                np.linalg.norm(acceleration, axis=1).max()], rtol=1e-6)
    velocity[0, 0] = np.nan
    sim.state_0.particle_qd.assign(velocity)
    with pytest.raises(RuntimeError, match="Non-finite velocity"):
        sim._cfl_maxima()
    velocity[0, 0] = 0.0
    sim.state_0.particle_qd.assign(velocity)
    acceleration[0, 0] = np.inf
    sim.solver._particle_accel.assign(acceleration)
    with pytest.raises(RuntimeError, match="Non-finite WCSPH acceleration"):
        sim._cfl_maxima()


def test_viscous_timestep_damps_velocity_and_completes_frame():
    """unit_test_dam_break.py:132-151"""
    sim = DamBreakSimulation(make_scene(512), "wcsph", device="cuda:0", wcsph_viscosity=10.0)
    velocity = np.zeros((sim.scene.count, 3), dtype=np.float32)
    velocity[:, 0] = np.where(np.arange(sim.scene.count) % 2, -0.01, 0.01)
    sim.state_0.particle_qd.assign(velocity)
    sim.model.set_gravity((0.0, 0.0, 0.0))
    dts = []
    advance = sim._substep

    def record_and_advance(dt):
        dts.append(dt)
        advance(dt)

    sim._substep = record_and_advance
    sim.step()
    sim.check_finite()
    assert max(dts) <= 0.025 * sim.solver.smoothing_length ** 2 / 10.0
    assert sum(dts) == pytest.approx(sim.frame_dt)
    assert np.sum(sim.state_0.particle_qd.numpy() ** 2) < np.sum(velocity ** 2)


@pytest.mark.parametrize("method", list(SOLVERS))
def test_two_frames_track_the_oracle(method, oracle):
    """Two display frames of the 540-particle dam break: GPU driver vs the
    oracle stepping with the same dt sequence."""
    scene = make_scene(512)
    sim = DamBreakSimulation(scene, method, device="cuda:0")
    dts = []
    advance = sim._substep

    def record(dt):
        dts.append(dt)
        advance(dt)

    sim._substep = record
    om = oracle_model(oracle, sim.model, method)
    ocfg = oracle_config(oracle, sim.solver.config)
    ost = oracle.OState(sim.state_0.particle_q.numpy(), sim.state_0.particle_qd.numpy(),
                        {k: getattr(sim.state_0.sph, k).numpy() for k in sim.state_0.sph.names()})
    sim.step()
    sim.step()
    fn = {"wcsph": oracle.wcsph_step, "dfsph": oracle.dfsph_step, "pbf": oracle.pbf_step}[method]
    for dt in dts:
        ost, _ = fn(om, ost, ocfg, dt)
    ex = rel_err(sim.state_0.particle_q.numpy(), ost.x, 1e-2)
    ev = rel_err(sim.state_0.particle_qd.numpy(), ost.v, 1e-3)
    er = rel_err(sim.state_0.sph.rho.numpy(), ost.sph["rho"], 1.0)
    print(f"{method}: {len(dts)} substeps, rel err x={ex:.2e} v={ev:.2e} rho={er:.2e}")
    assert max(ex, ev, er) <= 1e-5


@pytest.mark.parametrize("method", ["wcsph", "pbf", "dfsph"])
def test_substep_replays_from_a_hip_graph(method):
    """A captured substep replays bit-identically to the eager substep, with
    eager work and synchronisation between replays."""
    sim = DamBreakSimulation(make_scene(10000), method, device="cuda:0")
    sim.step()
    dt = sim.frame_dt / 8
    s_in, s_out = sim.state_0, sim.state_1
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        sim.solver.step(s_in, s_out, None, None, dt)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    ref_x, ref_v = s_out.particle_q.numpy(), s_out.particle_qd.numpy()
    ref_rho = s_out.sph.rho.numpy()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        sim.solver.step(s_in, s_out, None, None, dt)
    for rep in range(3):
        s_out.particle_q.zero_()
        s_out.sph.rho.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(s_out.particle_q.numpy(), ref_x, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(s_out.particle_qd.numpy(), ref_v, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(s_out.sph.rho.numpy(), ref_rho, err_msg=f"replay {rep}")
