"""The cooperative cell scan (fs_coop.h) against the per-lane scan.

The cooperative variants of the WCSPH density pass stage each query group's
candidates in LDS, test them against every particle of the group at once with
a widened superset test, and re-test the passing ones with the per-lane
scan's own prefilter.  So every output -- density, pressure, the substep's
positions and velocities, and the recorded neighbour lists entry by entry --
must equal the per-lane scan's bit for bit (the per-lane scan is the one the
solver parity tests compare with the oracle).  The cases cover the paths the
cooperative scan has: ordinary groups, inactive and masked particles, dense
clumps whose groups overflow the LDS chunks (the particles then scan alone),
aliased buckets, cell coordinates at and below -2^20 (queries that revisit
bucket 0), a non-finite position, and the batched form; the lists are also
checked against the oracle's within-h sets."""
import numpy as np
import pytest
import torch

from fluid_simulation_amd import autotune as at
from tests.cases import build_model, make_solver
from tests.test_gpu_lists import _check_lists, _recorded
from tests.test_gpu_solvers import _compressed_case

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset_variants():
    yield
    for p in at.PASSES:
        at.select(p, 0)


def _coop_variants(name):
    out = [v for v, sh in enumerate(at.describe(name)) if sh["coop"] > 0]
    assert out, f"{name} has no cooperative variant"
    return out


def _per_lane_variant(name):
    return next(v for v, sh in enumerate(at.describe(name)) if sh["coop"] == 0)


def _wcsph_case(kind, seed=3):
    """(model, cfg, dt) of a WCSPH case."""
    if kind == "compressed":
        model, cfg = _compressed_case("wcsph", seed=seed, surface_tension=0.01)
        return model, cfg, 2.0e-4
    _, cfg = _compressed_case("wcsph", seed=seed, inactive=False)
    d = 0.02
    rng = np.random.default_rng(seed)
    if kind == "clumps":
        # Gaussian clumps far denser than the fluid: a query group then has
        # thousands of candidates and overflows the tile's LDS chunks
        c = rng.uniform(-0.1, 0.1, (3, 3))
        pts = c[rng.integers(0, 3, 6000)] + rng.normal(0.0, 0.004, (6000, 3))
    elif kind == "aliasing":
        # 140 cells of h along x: buckets of cells 128 apart coincide
        h = 2.0 * d
        x = rng.uniform(-70 * h, 70 * h, (5000, 1))
        yz = rng.uniform(-0.04, 0.04, (5000, 2))
        pts = np.hstack([x, yz])
    elif kind == "far_clamped":
        # cell coordinates at and below -2^20 hash to bucket 0 on their axis,
        # which the query then visits once per such cell
        h = 2.0 * d
        edge = -(2 ** 20) * h
        near = np.column_stack([rng.uniform(edge - 0.2, edge + 0.2, 1500), rng.uniform(-0.05, 0.05, (1500, 2))])
        far_y = np.column_stack([rng.uniform(-0.05, 0.05, 800), rng.uniform(edge - 0.2, edge + 0.2, 800),
                                 rng.uniform(-0.05, 0.05, 800)])
        lattice = rng.uniform(-0.08, 0.08, (2000, 3))
        pts = np.vstack([near, far_y, lattice])
    elif kind == "nonfinite":
        pts = rng.uniform(-0.1, 0.1, (4000, 3))
        pts[17] = (np.nan, 0.0, 0.0)
        pts[2001] = (0.01, np.inf, 0.0)
    else:
        raise KeyError(kind)
    pts = pts.astype(np.float32)
    n = pts.shape[0]
    vel = (rng.normal(size=(n, 3)) * 0.2).astype(np.float32)
    flags = np.ones(n, np.int32)
    flags[rng.choice(n, n // 20, replace=False)] = 0
    mat = {"rest_density": 1000.0, "surface_tension": 0.0, "viscosity": 1.0e-3, "stiffness": 250000.0,
           "exponent": 7.0}
    model = build_model("wcsph", pts, vel, mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat, flags=flags)
    return model, cfg, 1.0e-4


def _outputs(model, cfg, dt):
    solver = make_solver("wcsph", model, cfg)
    s_in, s_out = model.state(), model.state()
    solver.step(s_in, s_out, None, None, dt)
    torch.cuda.synchronize()
    n = model.particle_count
    cap, _pos, _flags, cnt, nbr = _recorded(solver, n)
    live = np.arange(cap)[:, None] < np.minimum(cnt, cap)[None, :]
    lists = np.where(live, nbr, -1)
    return solver, [s_out.particle_q.numpy(), s_out.particle_qd.numpy(), s_out.sph.rho.numpy(),
                    s_out.sph.pressure.numpy(), cnt, lists]


NAMES = ["x", "v", "rho", "pressure", "list counts", "list entries"]


@pytest.mark.parametrize("kind", ["compressed", "clumps", "aliasing", "far_clamped", "nonfinite"])
def test_coop_density_equals_per_lane_scan(kind):
    model, cfg, dt = _wcsph_case(kind)
    at.select("wcsph_density", _per_lane_variant("wcsph_density"))
    _, ref = _outputs(model, cfg, dt)
    for v in _coop_variants("wcsph_density"):
        sh = at.describe("wcsph_density")[v]
        print(f"{kind}: wcsph_density variant {v} {sh}", flush=True)
        at.select("wcsph_density", v)
        _, got = _outputs(model, cfg, dt)
        for name, a, b in zip(NAMES, got, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"{kind} variant {v}: {name}")


def test_coop_lists_are_the_oracle_within_h_sets(oracle):
    model, cfg, dt = _wcsph_case("compressed")
    for v in _coop_variants("wcsph_density"):
        at.select("wcsph_density", v)
        solver, _ = _outputs(model, cfg, dt)
        checked = _check_lists(solver, model, solver.smoothing_length, oracle)
        print(f"coop variant {v}: {checked} lists match", flush=True)
        assert checked > 0.8 * model.particle_count


def test_coop_batched_density_equals_per_lane_scan():
    from fluid_simulation_amd.batched import BatchedWCSPH, make_batch_scenes

    def run(variant):
        at.select("batch_density", variant)
        batch = BatchedWCSPH(make_batch_scenes(5, seed=4, count_range=(1500, 7000)), device="cuda:0")
        batch.begin_frame()
        for _ in range(4):
            batch.substep()
        torch.cuda.synchronize()
        return batch.positions.cpu().numpy(), batch.velocities.cpu().numpy()

    ref = run(_per_lane_variant("batch_density"))
    for v in _coop_variants("batch_density"):
        print(f"batch_density variant {v}: {at.describe('batch_density')[v]}", flush=True)
        for a, b in zip(run(v), ref):
            np.testing.assert_array_equal(a, b)
