"""Neighbour lists that grow when they overflow.

A solver built with its class's default list capacity records, on the
device, the longest list an overflowing particle needed; the next step reads
that (copied to pinned memory after the step, no synchronisation) and
re-initialises the workspace with longer lists.  Lists only save rescans, so
every result must equal a solver's without lists, bit for bit, before and
after the growth.  An explicit capacity stays fixed."""
import numpy as np
import pytest
import torch

from tests.cases import build_model, make_solver
from tests.test_gpu_solvers import _compressed_case

pytestmark = pytest.mark.gpu


def _clumped_model(method, seed=5):
    """Dense Gaussian clumps: several hundred neighbours within h."""
    _, cfg = _compressed_case(method, seed=seed, inactive=False)
    d = 0.02
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.08, 0.08, (2, 3))
    pts = (c[rng.integers(0, 2, 4000)] + rng.normal(0.0, 0.006, (4000, 3))).astype(np.float32)
    vel = (rng.normal(size=pts.shape) * 0.1).astype(np.float32)
    mat = {"rest_density": 1000.0, "surface_tension": 0.0, "viscosity": 1.0e-3}
    if method == "wcsph":
        mat.update(stiffness=250000.0, exponent=7.0)
    return build_model(method, pts, vel, mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat), cfg


def _step(solver, model, dt):
    s_in, s_out = model.state(), model.state()
    solver.step(s_in, s_out, None, None, dt)
    torch.cuda.synchronize()
    out = [s_out.particle_q.numpy(), s_out.particle_qd.numpy(), s_out.sph.rho.numpy()]
    if hasattr(s_out.sph, "factor"):
        out.append(s_out.sph.factor.numpy())
    return out


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_lists_grow_after_an_overflow_with_unchanged_results(method):
    model, cfg = _clumped_model(method)
    dt = 1.0e-4 if method == "wcsph" else 1.0e-3
    ref = _step(make_solver(method, model, cfg, neighbor_list_capacity=0), model, dt)
    solver = make_solver(method, model, cfg)
    cap0 = solver.neighbor_list_capacity
    assert solver.adaptive_lists
    got = _step(solver, model, dt)
    need = solver.list_need()
    print(f"{method}: default capacity {cap0}, longest overflowed list {need}", flush=True)
    assert need > cap0
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    got = _step(solver, model, dt)      # reads the overflow word queued by the first step, then grows
    assert solver.neighbor_list_capacity >= need
    assert solver.list_need() == 0
    print(f"{method}: grown to {solver.neighbor_list_capacity}", flush=True)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_explicit_capacity_stays_fixed_and_adapt_lists_grows_on_request():
    model, cfg = _clumped_model("wcsph")
    fixed = make_solver("wcsph", model, cfg, neighbor_list_capacity=64)
    assert not fixed.adaptive_lists
    for _ in range(3):
        _step(fixed, model, 1.0e-4)
    assert fixed.neighbor_list_capacity == 64 and fixed.list_need() > 64
    assert not fixed.adapt_lists()
    grown = make_solver("wcsph", model, cfg)
    _step(grown, model, 1.0e-4)
    assert grown.adapt_lists()
    assert grown.neighbor_list_capacity > 96
    _step(grown, model, 1.0e-4)
    assert not grown.adapt_lists()
