"""Batched WCSPH on the GPU: every scene of a batch evolves bit-identically to
a plain solve of that scene with the same dt sequence, and the device CFL dt
equals the dam-break driver's host formula (simulation.py:166-187)."""
import numpy as np
import pytest
import torch

from fluid_simulation_amd.batched import BatchedWCSPH, make_batch_scenes
from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

pytestmark = pytest.mark.gpu


def _plain(scene):
    return DamBreakSimulation(scene, "wcsph", device="cuda:0")


def _compare(batch, sims, what):
    x, v = batch.positions.cpu().numpy(), batch.velocities.cpu().numpy()
    rho, p = batch.rho.cpu().numpy(), batch.pressure.cpu().numpy()
    for k, sim in enumerate(sims):
        sl = batch.scene_slice(k)
        msg = f"{what}: scene {k} ({sim.scene.count} particles)"
        np.testing.assert_array_equal(x[sl], sim.state_0.particle_q.numpy(), err_msg=msg)
        np.testing.assert_array_equal(v[sl], sim.state_0.particle_qd.numpy(), err_msg=msg)
        np.testing.assert_array_equal(rho[sl], sim.state_0.sph.rho.numpy(), err_msg=msg)
        np.testing.assert_array_equal(p[sl], sim.state_0.sph.pressure.numpy(), err_msg=msg)


def test_batched_substeps_match_plain_solves_bitwise():
    scenes = make_batch_scenes(6, seed=11, count_range=(800, 12000))
    batch = BatchedWCSPH(scenes, device="cuda:0")
    sims = [_plain(s) for s in scenes]
    remaining = [sim.frame_dt for sim in sims]
    batch.begin_frame()
    for step in range(4):
        batch.substep()
        dts = batch.dts()
        batch.check_status()
        for k, sim in enumerate(sims):
            # the host policy in float64 with the plain solver's float64 h;
            # the substep takes it as float32
            expect = min(remaining[k], sim.cfl_limit(*sim._cfl_maxima()))
            assert dts[k] == np.float32(expect), f"scene {k} step {step}"
            remaining[k] -= expect
            sim._substep(expect)
        _compare(batch, sims, f"substep {step}")


def test_scene_dt_zero_freezes_a_scene_and_others_still_match():
    scenes = make_batch_scenes(3, seed=5, count_range=(1000, 6000))
    batch = BatchedWCSPH(scenes, device="cuda:0")
    sims = [_plain(s) for s in scenes]
    dt = np.array([1e-4, 0.0, 2e-4], np.float32)
    for step in range(3):
        batch.set_dt(dt)
        batch.advance()
        for k, sim in enumerate(sims):
            if dt[k] > 0:
                sim._substep(float(dt[k]))
    x = batch.positions.cpu().numpy()
    np.testing.assert_array_equal(x[batch.scene_slice(1)], scenes[1].positions)
    for k in (0, 2):
        sl = batch.scene_slice(k)
        np.testing.assert_array_equal(x[sl], sims[k].state_0.particle_q.numpy())
        np.testing.assert_array_equal(batch.velocities.cpu().numpy()[sl], sims[k].state_0.particle_qd.numpy())


def test_batched_frame_completes_every_scene_and_replays_from_a_graph():
    """A whole frame of per-scene CFL substeps; the substep sequence also
    replays from a HIP graph with eager work between replays."""
    scenes = make_batch_scenes(4, seed=7, count_range=(1500, 8000))
    batch = BatchedWCSPH(scenes, device="cuda:0")
    taken = batch.step_frame(chunk=8)
    assert taken >= 4
    assert batch.frame_done()
    assert np.isfinite(batch.positions.cpu().numpy()).all()
    # eager reference of two further substeps from the current state
    x0 = [t.clone() for t in batch.x]
    v0 = [t.clone() for t in batch.v]
    cur0, tot0 = batch.cur, batch.total_substeps
    batch.begin_frame()
    batch.substep()
    batch.substep()
    torch.cuda.synchronize()
    ref = (batch.positions.clone(), batch.velocities.clone(), batch.rho.clone())

    def restore():
        for a, b in zip(batch.x, x0):
            a.copy_(b)
        for a, b in zip(batch.v, v0):
            a.copy_(b)
        batch.cur, batch.total_substeps = cur0, tot0
        batch.remaining.fill_(batch.frame_dt)

    restore()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            batch.substep()
            batch.substep()
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(3):
        restore()
        batch.rho.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(batch.x[cur0].cpu().numpy(), ref[0].cpu().numpy(), err_msg=f"replay {rep}")
        np.testing.assert_array_equal(batch.v[cur0].cpu().numpy(), ref[1].cpu().numpy(), err_msg=f"replay {rep}")
        np.testing.assert_array_equal(batch.rho.cpu().numpy(), ref[2].cpu().numpy(), err_msg=f"replay {rep}")
