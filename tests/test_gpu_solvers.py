"""One-substep parity of the HIP solvers with the FP32 oracle.

Each case feeds the same input state to ``Solver*.step`` (GPU, through the
C-ABI) and to the oracle's restatement of the reference ``step``; integer
outputs must be identical and float outputs within 1e-5 relative (row-norm-wise
for vectors, with a stated floor).  The kernels evaluate the reference's
float32 expressions in the same order, so most outputs are bit-identical; the
exact-match fraction is printed.
"""
from dataclasses import replace

import numpy as np
import pytest
import torch

from fluid_simulation_amd.array import Array
from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation, solver_config, material
from fluid_simulation_amd.examples.dam_break.scene import make_scene
from tests.cases import build_model, lattice_block, make_solver, oracle_config, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1.0e-5


def _compare(name, gpu, ref, floor):
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    assert gpu.shape == ref.shape, name
    if ref.dtype.kind == "i":
        np.testing.assert_array_equal(gpu, ref, err_msg=name)
        return 1.0
    exact = float(np.mean(gpu.view(np.uint32) == ref.astype(np.float32).view(np.uint32)))
    err = rel_err(gpu, ref, floor)
    print(f"  {name:16s} exact={exact:.4f} rel_err={err:.3e}", flush=True)
    assert np.all(np.isfinite(gpu) == np.isfinite(ref)), name
    assert err <= TOL, f"{name}: rel err {err:.3e} > {TOL}"
    return exact


def _compressed_case(method, seed=0, inactive=True, surface_tension=0.0):
    """Jittered lattice at 0.9x spacing (positive pressure), random velocities,
    straddling two walls, with inactive and masked particles."""
    d = 0.02
    pts = lattice_block(12, 0.9 * d, origin=(-0.215, -0.1, -0.004), jitter=0.05, seed=seed)
    n = pts.shape[0]
    rng = np.random.default_rng(seed + 1)
    vel = (rng.normal(size=(n, 3)) * 0.5).astype(np.float32)
    flags = np.ones(n, np.int32)
    mask = np.zeros(n, np.int32)
    if inactive:
        flags[rng.choice(n, n // 10, replace=False)] = 0
        mask[rng.choice(n, n // 20, replace=False)] = 1
    mat = {"rest_density": 1000.0, "surface_tension": surface_tension, "viscosity": 1.0e-3}
    if method == "wcsph":
        mat.update(stiffness=250000.0, exponent=7.0)
    common = dict(particle_radius=0.5 * d, particle_length=d, smoothing_length_coff=2.0, boundary_damping=-0.3,
                  bound_width=0.2, bound_height=0.2, bound_length=0.15)
    from fluid_simulation_amd.solver.pbf import SolverPBF
    from fluid_simulation_amd.solver.sph.solver_dfsph import SolverDFSPH
    from fluid_simulation_amd.solver.sph.solver_wcsph import SolverWCSPH
    if method == "wcsph":
        cfg = SolverWCSPH.SphConfig(clamp_negative_pressure=True, **common)
    elif method == "dfsph":
        cfg = SolverDFSPH.SphConfig(max_rho_iterations=5, max_vel_iterations=6, warm_start=False, max_kappa=1.0,
                                    max_kappa_v=100.0, **common)
    else:
        cfg = SolverPBF.PbfConfig(max_iterations=4, lambda_epsilon=100.0, min_neighbors_for_lambda=8, scorr_k=0.0,
                                  max_delta_position=0.25 * d, xsph_c=0.05, **common)
    model = build_model(method, pts, vel, mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat, flags=flags, mask=mask)
    return model, cfg


def _run_pair(method, model, cfg, dt, oracle, state_in=None):
    solver = make_solver(method, model, cfg)
    s_in = model.state() if state_in is None else state_in
    s_out = model.state()
    solver.step(s_in, s_out, None, None, dt)
    torch.cuda.synchronize()
    om = oracle_model(oracle, model, method)
    ost = oracle.OState(s_in.particle_q.numpy(), s_in.particle_qd.numpy(),
                        {k: getattr(s_in.sph, k).numpy() for k in s_in.sph.names()})
    ocfg = oracle_config(oracle, cfg)
    fn = {"wcsph": oracle.wcsph_step, "dfsph": oracle.dfsph_step, "pbf": oracle.pbf_step}[method]
    oout, extras = fn(om, ost, ocfg, dt)
    return solver, s_out, oout, extras


def _check_outputs(method, solver, s_out, oout, extras):
    print(f"{method}:", flush=True)
    _compare("x", s_out.particle_q.numpy(), oout.x, 1.0e-2)
    _compare("v", s_out.particle_qd.numpy(), oout.v, 1.0e-3)
    _compare("rho", s_out.sph.rho.numpy(), oout.sph["rho"], 1.0)
    _compare("accel", solver._particle_accel.numpy(), extras["accel"], 1.0)
    if method == "wcsph":
        p = oout.sph["pressure"]
        _compare("pressure", s_out.sph.pressure.numpy(), p, max(1.0, 1e-3 * float(np.abs(p).max(initial=0))))
    if method == "dfsph":
        for k in ("factor", "rho_star", "d_rho_dt", "kappa", "kappa_v", "v_star", "neighbor_count"):
            ref = oout.sph[k]
            floor = 1e-3 if k == "v_star" else max(1e-12, 1e-3 * float(np.abs(ref).max(initial=0)))
            _compare(k, getattr(s_out.sph, k).numpy(), ref, floor)
        it = solver.last_iterations()
        assert it["density_iterations"] == extras["rho_iterations"]
        assert it["divergence_iterations"] == extras["vel_iterations"]
        assert it["density_corrections"] == extras.get("rho_corrections", 0)
        assert it["divergence_corrections"] == extras.get("vel_corrections", 0)
    if method == "pbf":
        for k in ("pbf_lambda", "neighbor_count", "delta_p"):
            ref = oout.sph[k]
            floor = max(1e-12, 1e-3 * float(np.abs(ref).max(initial=0)))
            _compare(k, getattr(s_out.sph, k).numpy(), ref, floor)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_compressed_block_one_substep(method, oracle):
    model, cfg = _compressed_case(method)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    solver, s_out, oout, extras = _run_pair(method, model, cfg, dt, oracle)
    _check_outputs(method, solver, s_out, oout, extras)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_coincident_and_nearly_coincident_particles(method, oracle):
    """Degenerate separations: exact duplicates of block particles (r = 0,
    where the kernel gradients return zero: r <= 1e-5 for the cubic spline,
    r <= 1e-6 for the spiky gradient), and a pair 1e-15 apart near the
    origin, whose squared distance (1e-30) is below the 2^-96 where the fast
    square root hands over to sqrtf.  Against the oracle, as the other
    one-substep cases."""
    model0, cfg = _compressed_case(method, seed=4, inactive=False)
    d = 0.02
    base = model0.particle_q.numpy()
    rng = np.random.default_rng(4)
    dup = base[rng.choice(base.shape[0], 6, replace=False)]
    tiny = np.array([[1.0e-15, 1.0e-15, 1.0e-15], [2.0e-15, 1.0e-15, 1.0e-15]], np.float32)
    pts = np.vstack([base, dup, tiny]).astype(np.float32)
    n = pts.shape[0]
    vel = (rng.normal(size=(n, 3)) * 0.5).astype(np.float32)
    mat = {"rest_density": 1000.0, "surface_tension": 0.01, "viscosity": 1.0e-3}
    if method == "wcsph":
        mat.update(stiffness=250000.0, exponent=7.0)
    model = build_model(method, pts, vel, mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    solver, s_out, oout, extras = _run_pair(method, model, cfg, dt, oracle)
    _check_outputs(method, solver, s_out, oout, extras)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_block_far_from_the_origin(method, oracle):
    """The compressed block moved to x = +1000 m, y = -700 m (bounds widened
    to hold it): coordinates whose float32 spacing is ~6e-5 m, so every
    separation carries rounding from the positions, and cell coordinates in
    the tens of thousands, which the 128-bucket hash wraps."""
    model0, cfg = _compressed_case(method, seed=5)
    d = 0.02
    pts = model0.particle_q.numpy() + np.array([1000.25, -700.5, 0.0], np.float32)
    vel = model0.particle_qd.numpy()
    flags = model0.particle_flags.numpy()
    mask = model0.sph.particle_mask.numpy()
    mat = {"rest_density": 1000.0, "surface_tension": 0.01, "viscosity": 1.0e-3}
    if method == "wcsph":
        mat.update(stiffness=250000.0, exponent=7.0)
    model = build_model(method, pts.astype(np.float32), vel, mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat,
                        flags=flags, mask=mask)
    cfg = replace(cfg, bound_width=2000.0, bound_height=2000.0)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    solver, s_out, oout, extras = _run_pair(method, model, cfg, dt, oracle)
    _check_outputs(method, solver, s_out, oout, extras)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_cohesion_one_substep(method, oracle):
    """Surface tension 0.01, the reference's default material value
    (solver_wcsph.py:348): the cohesion term ((gamma/m_i*m_j)*d)*W of
    cal_acc_with_non_pressure (solver_wcsph.py:78, solver_dfsph.py:82,
    solver_pbf.py:74) in every solver."""
    model, cfg = _compressed_case(method, seed=2, surface_tension=0.01)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    solver, s_out, oout, extras = _run_pair(method, model, cfg, dt, oracle)
    _check_outputs(method, solver, s_out, oout, extras)
    # the term is live: the oracle without it gives another acceleration
    model0, _ = _compressed_case(method, seed=2, surface_tension=0.0)
    _, _, _, extras0 = _run_pair(method, model0, cfg, dt, oracle)
    assert not np.array_equal(extras0["accel"], extras["accel"])


def test_cohesion_pairs_bitwise(oracle):
    """Isolated pairs, WCSPH without pressure or viscosity: each
    acceleration is one cohesion term, 0 - ((gamma/m_i*m_j)*d)*W.  The GPU's
    float32 result must equal the oracle's bit for bit; the other
    association, (gamma/m_i*m_j)*(d*W), gives other bits for some pairs."""
    d = 0.01
    h = 2 * d
    rng = np.random.default_rng(21)
    npairs = 400
    centres = np.stack(np.meshgrid(*(np.arange(8) * 3 * h,) * 3, indexing="ij"), -1).reshape(-1, 3)[:npairs]
    centres = centres + np.array([0.0, 0.0, 0.3])
    dirs = rng.normal(size=(npairs, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    r = rng.uniform(0.2, 0.95, npairs) * h
    pts = np.concatenate([centres, centres + dirs * r[:, None]]).astype(np.float32)
    n = len(pts)
    mat = {"rest_density": 1000.0, "surface_tension": 0.0123, "viscosity": 0.0, "stiffness": 0.0,
           "exponent": 7.0}
    from fluid_simulation_amd.solver.sph.solver_wcsph import SolverWCSPH
    cfg = SolverWCSPH.SphConfig(particle_radius=0.5 * d, particle_length=d, smoothing_length_coff=2.0,
                                boundary_damping=0.0, bound_width=5.0, bound_height=5.0, bound_length=5.0)
    model = build_model("wcsph", pts, np.zeros_like(pts), mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat,
                        gravity=(0.0, 0.0, 0.0))
    solver, s_out, oout, extras = _run_pair("wcsph", model, cfg, 1.0e-5, oracle)
    acc = solver._particle_accel.numpy()
    np.testing.assert_array_equal(acc.view(np.uint32), extras["accel"].view(np.uint32))
    assert np.abs(acc).max() > 0.0
    # the other association differs for some pairs
    f32 = np.float32
    g = f32(0.0123)
    m = f32(1000.0 * d ** 3)
    lib = oracle._lib("f32")
    partner = np.concatenate([np.arange(npairs) + npairs, np.arange(npairs)])
    other = np.zeros_like(acc)
    for i in range(n):
        dv = pts[i] - pts[partner[i]]
        dist = np.sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
        w = f32(lib.ora_cubic_w(lib._real(float(max(dist, f32(0.5 * d)))), lib._real(float(f32(h)))))
        other[i] = f32(0.0) - ((g / m) * m) * (dv * w)
    assert np.any(other != acc)


def test_dfsph_far_clamped_neighbor_counts(oracle):
    """Every particle active, positions whose query cells lie at or below
    -2^20 on an axis: those cells all hash to bucket 0, so a query visits the
    particle's own bucket more than once and the reference's count of active
    candidates j != i (solver_dfsph.py:150-160) skips it on every visit."""
    from tests.test_gpu_grid import _points
    from fluid_simulation_amd.solver.sph.solver_dfsph import SolverDFSPH

    pts, h = _points("far_clamped")
    pts = pts[np.all(np.abs(pts) < 1.0e6, axis=1)]     # the extreme points stay in the grid tests
    pts[:, 2] += 1.0                                     # above the floor: no bound clamps the step
    d = h / 2.0
    n = pts.shape[0]
    vel = np.zeros((n, 3), np.float32)
    mat = {"rest_density": 1000.0, "surface_tension": 0.0, "viscosity": 1.0e-3}
    cfg = SolverDFSPH.SphConfig(max_rho_iterations=3, max_vel_iterations=3, warm_start=False, max_kappa=1.0,
                                max_kappa_v=100.0, particle_radius=0.5 * d, particle_length=d,
                                smoothing_length_coff=2.0, boundary_damping=0.0, bound_width=1.0e5,
                                bound_height=1.0e5, bound_length=1.0e5)
    model = build_model("dfsph", pts, vel, mass=1000.0 * d ** 3, radius=0.5 * d, mat=mat)
    solver, s_out, oout, extras = _run_pair("dfsph", model, cfg, 1.0e-4, oracle)
    counts = s_out.sph.neighbor_count.numpy()
    np.testing.assert_array_equal(counts, oout.sph["neighbor_count"])
    # the far cells do repeat buckets: some count exceeds the particle's
    # distinct candidates minus one
    og = extras["grid2"]
    off, idx = og.candidates(oout.x, h)
    distinct = np.array([len(np.unique(idx[off[k]:off[k + 1]])) - 1 for k in range(n)])
    assert np.any(counts > distinct)
    _check_outputs("dfsph", solver, s_out, oout, extras)


@pytest.mark.parametrize("loop", ["density", "divergence"])
@pytest.mark.parametrize("exact", [False, True])
def test_dfsph_residual_at_eta(loop, exact, oracle):
    """The break test `iter > 0 and residual <= eta` with eta equal to the
    reference's float32 residual of iteration 1, and with eta one float64
    ulp below it: the loop must stop at iteration 1 in the first case and run
    on in the second, as numpy's float32 mean decides
    (solver_dfsph.py:579-596, 733-735, 849-851).  A float64 mean of the same
    values lies within float32 rounding of eta and would decide one of the
    two cases the other way."""
    model, cfg = _compressed_case("dfsph", seed=4)
    dt = 1.0e-3
    # the residuals of a run that never stops the loop under test (the
    # divergence loop's inputs depend on where the density loop stops, so
    # that one keeps its eta)
    probe = replace(cfg, eta_density=-1.0) if loop == "density" else replace(cfg, eta_divergence=-1.0)
    _, _, _, ex = _run_pair("dfsph", model, probe, dt, oracle)
    key = "rho_residuals" if loop == "density" else "vel_residuals"
    r1 = ex[key][1]
    assert r1 > 0.0
    for eta, stops in ((r1, True), (float(np.nextafter(r1, -np.inf)), False)):
        field = "eta_density" if loop == "density" else "eta_divergence"
        c = replace(cfg, **{field: eta})
        solver = make_solver("dfsph", model, c)
        solver.exact_residual = exact
        s_in, s_out = model.state(), model.state()
        solver.step(s_in, s_out, None, None, dt)
        torch.cuda.synchronize()
        om = oracle_model(oracle, model, "dfsph")
        ost = oracle.OState(s_in.particle_q.numpy(), s_in.particle_qd.numpy(),
                            {k: getattr(s_in.sph, k).numpy() for k in s_in.sph.names()})
        oout, extras = oracle.dfsph_step(om, ost, oracle_config(oracle, c), dt)
        it = solver.last_iterations()
        name = "density_iterations" if loop == "density" else "divergence_iterations"
        expect = extras["rho_iterations" if loop == "density" else "vel_iterations"]
        assert (expect == 2) == stops, (eta, expect)
        assert it[name] == expect, (loop, eta, it, expect)
        _check_outputs("dfsph", solver, s_out, oout, extras)


def test_dfsph_warm_start_one_substep(oracle):
    model, cfg = _compressed_case("dfsph", seed=5)
    cfg = replace(cfg, warm_start=True)
    s_in = model.state()
    rng = np.random.default_rng(11)
    s_in.sph.kappa.assign(rng.uniform(0, 0.5, model.particle_count).astype(np.float32))
    s_in.sph.kappa_v.assign(rng.uniform(0, 5.0, model.particle_count).astype(np.float32))
    solver, s_out, oout, extras = _run_pair("dfsph", model, cfg, 1.0e-3, oracle, state_in=s_in)
    _check_outputs("dfsph", solver, s_out, oout, extras)


@pytest.mark.parametrize("variant", ["rebuild", "scorr", "no_xsph", "zero_iterations"])
def test_pbf_config_variants(variant, oracle):
    model, cfg = _compressed_case("pbf", seed=9)
    cfg = {"rebuild": replace(cfg, rebuild_grid_each_iteration=True),
           "scorr": replace(cfg, scorr_k=1.0e-4, scorr_n=4.0, scorr_q=0.3, max_delta_position=0.0),
           "no_xsph": replace(cfg, xsph_c=0.0, boundary_epsilon=1e-4),
           "zero_iterations": replace(cfg, max_iterations=0)}[variant]
    solver, s_out, oout, extras = _run_pair("pbf", model, cfg, 1.0e-3, oracle)
    _check_outputs("pbf", solver, s_out, oout, extras)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_dam_break_one_substep_after_a_frame(method, oracle):
    """Advance the 9,936-particle dam break one frame on the GPU, then compare
    the one-substep map from that state."""
    scene = make_scene(10000)
    sim = DamBreakSimulation(scene, method, device="cuda:0")
    sim.step()
    dt = sim.frame_dt / 4 if method == "pbf" else min(sim.frame_dt / 4, sim.cfl_limit(*sim._cfl_maxima()))
    solver, s_out, oout, extras = _run_pair(method, sim.model, sim.solver.config, dt, oracle, state_in=sim.state_0)
    _check_outputs(method, solver, s_out, oout, extras)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_list_capacity_does_not_change_results(method):
    """Scanning passes record each particle's neighbours for the passes that
    follow; a particle whose list overflows scans the grid again.  Capacity 0
    (every pass scans), 8 (most lists overflow) and the default 96 must give
    bit-identical states: the neighbour visit order, and so every float sum,
    is the same on all three routes."""
    model, cfg = _compressed_case(method, seed=3)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    outs = []
    for cap in (0, 8, 96):
        solver = make_solver(method, model, cfg, neighbor_list_capacity=cap)
        assert solver.neighbor_list_capacity == cap
        s_in, s_out = model.state(), model.state()
        for _ in range(3):
            solver.step(s_in, s_out, None, None, dt)
            s_in, s_out = s_out, s_in
        torch.cuda.synchronize()
        outs.append((s_in.particle_q.numpy(), s_in.particle_qd.numpy(), s_in.sph.rho.numpy()))
    for k, name in enumerate(("x", "v", "rho")):
        ref = outs[2][k].view(np.uint32)
        for cap, o in zip((0, 8), outs[:2]):
            assert np.array_equal(o[k].view(np.uint32), ref), f"{method} {name}: capacity {cap} differs"


@pytest.mark.parametrize("method,pass_name", [("wcsph", "wcsph_force"), ("dfsph", "dfsph_replay"),
                                              ("pbf", "pbf_replay"), ("wcsph", "wcsph_density"),
                                              ("dfsph", "dfsph_scan"), ("pbf", "pbf_scan")])
def test_lane_groups_with_overflowing_lists(method, pass_name):
    """Passes under 2 and 4 lanes per particle.  Replaying passes: the
    group's lanes evaluate alternate list entries and fold them in list
    order; a particle whose list overflowed is scanned by the group's first
    lane alone.  Scanning passes: the group tests consecutive candidates,
    compacts the survivors in visit order, records them and folds their
    terms in that order.  With capacity 0 (every pass scans), 8 (most lists
    overflow) and 96, every lanes-per-particle variant matches a one-lane
    variant bit for bit."""
    from fluid_simulation_amd import autotune as at

    model, cfg = _compressed_case(method, seed=3, surface_tension=0.5 if method == "wcsph" else 0.0)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    shapes = at.describe(pass_name)
    picks = [v for v, sh in enumerate(shapes) if sh["lanes_per_particle"] > 1]
    ref = next(v for v, sh in enumerate(shapes) if sh["lanes_per_particle"] == 1)
    assert {shapes[v]["lanes_per_particle"] for v in picks} == {2, 4}
    extra = {"wcsph": ("pressure",), "dfsph": ("kappa", "kappa_v"), "pbf": ("pbf_lambda",)}[method]
    try:
        for cap in (0, 8, 96):
            outs = {}
            for v in [ref] + picks:
                at.select(pass_name, v)
                solver = make_solver(method, model, cfg, neighbor_list_capacity=cap)
                s_in, s_out = model.state(), model.state()
                for _ in range(3):
                    solver.step(s_in, s_out, None, None, dt)
                    s_in, s_out = s_out, s_in
                torch.cuda.synchronize()
                outs[v] = [s_in.particle_q.numpy(), s_in.particle_qd.numpy(), s_in.sph.rho.numpy()] + \
                          [getattr(s_in.sph, name).numpy() for name in extra]
            for v in picks:
                for k, (a, b) in enumerate(zip(outs[v], outs[ref])):
                    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
                        f"{method} capacity {cap}, variant {shapes[v]}: output {k} differs"
    finally:
        at.select(pass_name, 0)


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_fixup_gather_without_packed_records(method):
    """The grid build packs mass, rest density, flags and mask of each particle
    into one record before its fix-up gathers them in sorted order; a
    workspace without that buffer (packed = NULL) gathers the four arrays one
    by one.  Both routes must give bit-identical states."""
    model, cfg = _compressed_case(method, seed=5)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    outs = []
    for packed in (True, False):
        solver = make_solver(method, model, cfg)
        if not packed:
            solver._work.packed = None
        s_in, s_out = model.state(), model.state()
        for _ in range(3):
            solver.step(s_in, s_out, None, None, dt)
            s_in, s_out = s_out, s_in
        torch.cuda.synchronize()
        outs.append((s_in.particle_q.numpy(), s_in.particle_qd.numpy(), s_in.sph.rho.numpy()))
    for k, name in enumerate(("x", "v", "rho")):
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), f"{method} {name}"


def test_neighbour_lists_past_2_31_entries():
    """Lists whose entry index (entry * particles + slot) passes 2^31: 1M
    uniform particles plus a ball of 2,400 particles well inside one smoothing
    length, so each ball particle records more than 2^31 / n entries.  The
    states after one WCSPH substep must be bit-identical with and without
    lists (capacity 0 scans every pass)."""
    from fluid_simulation_amd.examples import microbench as mb
    from fluid_simulation_amd.solver.sph.solver_wcsph import SolverWCSPH

    n_ball = 2400
    n = 1_000_000
    h = mb.smoothing_length(n)
    rng = np.random.default_rng(17)
    pts = mb.uniform_points(n - n_ball)
    ball = np.array([0.5, 0.5, 0.5]) + rng.uniform(-0.15 * h, 0.15 * h, (n_ball, 3))
    pts = np.concatenate([pts, ball.astype(np.float32)])
    model, cfg = mb.wcsph_model(pts)
    cap = 2560              # the ball's ~2,440 neighbours fit
    assert cap * n > 2 ** 31 and n_ball - 1 > 2 ** 31 // n
    outs = []
    for c in (0, cap):
        solver = SolverWCSPH(model, cfg, neighbor_list_capacity=c)
        s0, s1 = model.state(), model.state()
        solver.step(s0, s1, None, None, 1.0e-6)
        torch.cuda.synchronize()
        outs.append([s1.particle_q.numpy(), s1.particle_qd.numpy(), s1.sph.rho.numpy()])
        del solver, s0, s1
        torch.cuda.empty_cache()
    for k, name in enumerate(("x", "v", "rho")):
        assert np.all(np.isfinite(outs[1][k])), name
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), name
    assert float(outs[1][2][-n_ball:].min()) > 50.0 * float(np.median(outs[1][2][: n - n_ball]))
