"""Launch variants of the tunable passes give bit-identical results, and the
autotuner selects one and leaves the simulation state untouched."""
import numpy as np
import pytest
import torch

from fluid_simulation_amd import autotune as at
from fluid_simulation_amd.batched import BatchedWCSPH, make_batch_scenes
from fluid_simulation_amd.examples.dam_break.scene import make_scene
from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset_variants():
    yield
    for p in at.PASSES:
        at.select(p, 0)


def _plain_substep(sim, dt):
    s_out = sim.model.state()
    sim.solver.step(sim.state_0, s_out, None, None, dt)
    torch.cuda.synchronize()
    return s_out.particle_q.numpy(), s_out.particle_qd.numpy(), s_out.sph.rho.numpy()


def test_plain_variants_are_bit_identical():
    sim = DamBreakSimulation(make_scene(10000), "wcsph", device="cuda:0")
    sim.step()
    ref = _plain_substep(sim, 1e-4)
    for name in ("wcsph_density", "wcsph_force"):
        for v, (block, depth) in enumerate(at.variants(name)):
            print(f"{name} variant {v}: block {block} depth {depth}", flush=True)
            at.select(name, v)
            got = _plain_substep(sim, 1e-4)
            for a, b in zip(got, ref):
                np.testing.assert_array_equal(a, b)
        at.select(name, 0)


@pytest.mark.parametrize("method", ["dfsph", "pbf"])
def test_dfsph_pbf_variants_are_bit_identical(method):
    """The scanning and list-replaying passes of DFSPH and PBF under every
    launch variant (workgroup, LDS list depth, replay unroll): the same
    substep, bit for bit, and the same DFSPH iteration counts."""
    sim = DamBreakSimulation(make_scene(10000), method, device="cuda:0")
    sim.step()
    dt = sim.frame_dt / 4

    def run():
        s_out = sim.model.state()
        sim.solver.step(sim.state_0, s_out, None, None, dt)
        torch.cuda.synchronize()
        out = [s_out.particle_q.numpy(), s_out.particle_qd.numpy(), s_out.sph.rho.numpy()]
        if method == "dfsph":
            out += [s_out.sph.factor.numpy(), s_out.sph.kappa_v.numpy(),
                    np.array(list(sim.solver.last_iterations().values()))]
        else:
            out += [s_out.sph.pbf_lambda.numpy(), s_out.sph.neighbor_count.numpy()]
        return out

    ref = run()
    for name in at.SOLVER_PASSES[method]:
        for v, shape in enumerate(at.describe(name)):
            print(f"{name} variant {v}: {shape}", flush=True)
            at.select(name, v)
            for a, b in zip(run(), ref):
                np.testing.assert_array_equal(a, b)
        at.select(name, 0)


def test_batched_variants_are_bit_identical_and_autotune_restores_state():
    batch = BatchedWCSPH(make_batch_scenes(4, seed=2, count_range=(2000, 9000)), device="cuda:0")
    batch.begin_frame()
    for _ in range(3):
        batch.substep()
    x0 = batch.positions.cpu().numpy()
    res = batch.autotune(use_cache=False)
    assert set(res) == {"batch_density", "batch_force"}
    np.testing.assert_array_equal(batch.positions.cpu().numpy(), x0)
    dts = batch.dts()
    outs = []
    for name in ("batch_density", "batch_force"):
        for v in range(len(at.variants(name))):
            at.select(name, v)
            cur = batch.cur
            xs, vs = batch.x[cur].clone(), batch.v[cur].clone()
            batch.set_dt(dts)
            batch.advance()
            outs.append((batch.positions.cpu().numpy(), batch.velocities.cpu().numpy(), batch.rho.cpu().numpy()))
            batch.cur = cur
            batch.x[cur].copy_(xs)
            batch.v[cur].copy_(vs)
        at.select(name, 0)
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            np.testing.assert_array_equal(a, b)
