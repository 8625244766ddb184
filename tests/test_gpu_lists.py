"""The neighbour lists the scanning passes record, read straight from the
solver's workspace and compared with the oracle's within-h candidate sets.

The hot path's scan (buffer-resource batch loads that read past a row's end,
LDS compaction, one lane or a group of 2 or 4 lanes per particle) records,
per particle, the active candidates that pass its within-cut prefilter, in
Warp's visit order.  For every particle, the entries at q <= 1 (the same
float32 test as the oracle's, r / h <= 1 with r = |x_i - x_j|) must equal the
oracle's within-h candidates (`ora_grid_candidates`, mode 0) that are
active, in the same order; the remaining entries may only be the prefilter's
slack band (|r|^2 <= 1.0001 h^2); and no active within-h neighbour may be
missing.  The solvers' outputs already depend on these lists bit for bit;
this checks them directly, for every lanes-per-particle variant of each
solver's scanning pass."""
import numpy as np
import pytest
import torch

from fluid_simulation_amd import _lib
from fluid_simulation_amd import autotune as at
from tests.cases import make_solver
from tests.test_gpu_solvers import _compressed_case

pytestmark = pytest.mark.gpu

SLACK = np.float32(1.0001)


def _work_view(solver, field, dtype, count):
    mem = solver._work_mem
    off = int(getattr(solver._work, field)) - mem.data_ptr()
    size = torch.tensor([], dtype=dtype).element_size()
    assert 0 <= off and off + count * size <= mem.numel()
    return mem[off:off + count * size].view(dtype)


def _recorded(solver, n):
    cap = int(solver.neighbor_list_capacity)
    pos = _work_view(solver, "pos", torch.float32, 4 * n).reshape(n, 4).cpu().numpy()
    flags = _work_view(solver, "flags", torch.int32, n).cpu().numpy()
    cnt = _work_view(solver, "nbr_count", torch.int32, n).cpu().numpy()
    nbr = _work_view(solver, "nbr", torch.int32, cap * n).reshape(cap, n).cpu().numpy()
    return cap, pos, flags, cnt, nbr


def _check_lists(solver, model, h, oracle):
    n = model.particle_count
    grid = model.particle_grid
    pid = grid.point_ids.cpu().numpy()
    cap, pos, flags, cnt, nbr = _recorded(solver, n)
    x = np.empty((n, 3), np.float32)
    x[pid] = pos[:, :3]
    active = np.zeros(n, bool)
    active[pid] = (flags & 1) != 0
    cnt_orig = np.empty(n, np.int64)
    cnt_orig[pid] = cnt
    assert np.all(cnt_orig[~active] == 0)
    overflow = cnt_orig > cap
    assert overflow.mean() < 0.05, f"{overflow.sum()} of {n} lists overflowed"
    og = oracle.OracleGrid()
    og.build(x, h)
    off, idx = og.candidates(x, h, within="cubic")
    h32 = np.float32(h)
    checked = 0
    for s in range(n):
        i = int(pid[s])
        if not active[i] or overflow[i]:
            continue
        got = pid[nbr[:cnt[s], s]]
        d = x[i][None, :] - x[got]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        within = (np.sqrt(r2) / h32) <= np.float32(1.0)
        want = idx[off[i]:off[i + 1]]
        want = want[active[want]]
        np.testing.assert_array_equal(got[within], want, err_msg=f"particle {i}: within-h entries")
        assert np.all(r2[~within] <= h32 * h32 * SLACK), f"particle {i}: entry beyond the slack band"
        assert np.all(active[got]), f"particle {i}: inactive neighbour recorded"
        checked += 1
    return checked


def _scan(method, solver, s_in, s_out, dt):
    """Run the solver up to a list-recording scan whose positions are the
    ones the grid was built from: the whole substep for WCSPH (density pass)
    and DFSPH (density/factor after the second build); for PBF, whose
    iterations scan moved positions over the grid built at the predicted
    ones, the predict pass and the first density pass only, through the
    per-pass entry points with the arguments the whole step would get."""
    if method != "pbf":
        solver.step(s_in, s_out, None, None, dt)
        return
    lib = _lib.load()
    captured = {}

    def grab(g, w, m, p, io, s):
        captured["args"] = (g, w, m, p, io, s)
        return 0

    whole = lib.fs_pbf_step
    lib.fs_pbf_step = grab
    try:
        solver.step(s_in, s_out, None, None, dt)
    finally:
        lib.fs_pbf_step = whole
    g, w, m, p, io, s = captured["args"]
    _lib.check(lib.fs_pbf_predict(g, w, m, p, io, s), "fs_pbf_predict")
    _lib.check(lib.fs_pbf_density(g, w, m, p, io, s), "fs_pbf_density")


@pytest.mark.parametrize("method,pass_name", [("wcsph", "wcsph_density"), ("dfsph", "dfsph_scan"),
                                              ("pbf", "pbf_scan")])
def test_recorded_lists_are_the_oracle_within_h_sets(method, pass_name, oracle):
    model, cfg = _compressed_case(method, seed=6)
    dt = 2.0e-4 if method == "wcsph" else 1.0e-3
    shapes = at.describe(pass_name)
    picks = {}
    for v, sh in enumerate(shapes):
        picks.setdefault(sh["lanes_per_particle"], v)
    assert set(picks) == {1, 2, 4}
    try:
        for lpp, v in sorted(picks.items()):
            at.select(pass_name, v)
            solver = make_solver(method, model, cfg)
            s_in, s_out = model.state(), model.state()
            _scan(method, solver, s_in, s_out, dt)
            torch.cuda.synchronize()
            checked = _check_lists(solver, model, solver.smoothing_length, oracle)
            print(f"{method} {pass_name} lanes per particle {lpp}: {checked} lists match", flush=True)
            assert checked > 0.8 * model.particle_count * 0.9
    finally:
        at.select(pass_name, 0)


def _check_lists_vectorized(solver, model, h, oracle):
    """_check_lists for large counts: the same three checks over all
    particles at once."""
    n = model.particle_count
    pid = model.particle_grid.point_ids.cpu().numpy().astype(np.int64)
    cap, pos, flags, cnt, nbr = _recorded(solver, n)
    inv = np.empty(n, np.int64)
    inv[pid] = np.arange(n)
    x = np.empty((n, 3), np.float32)
    x[pid] = pos[:, :3]
    active = np.zeros(n, bool)
    active[pid] = (flags & 1) != 0
    c = cnt[inv].astype(np.int64)                       # by original index
    assert np.all(c[~active] == 0)
    ok = active & (c <= cap)
    assert (~ok & active).mean() < 0.05
    ent = pid[nbr.T[inv]]                                # (n, cap) original indices
    live = (np.arange(cap)[None, :] < np.minimum(c, cap)[:, None]) & ok[:, None]
    d = x[:, None, :] - x[np.where(live, ent, 0)]
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    within = (np.sqrt(r2) / np.float32(h)) <= np.float32(1.0)
    assert np.all(r2[live & ~within] <= np.float32(h) * np.float32(h) * SLACK)
    assert np.all(active[ent[live]])
    got = ent[live & within]
    got_n = (live & within).sum(axis=1)
    og = oracle.OracleGrid()
    og.build(x, h)
    off, idx = og.candidates(x, h, within="cubic")
    row = np.repeat(np.arange(n), np.diff(off))
    keep = active[idx] & ok[row]
    want = idx[keep]
    want_n = np.bincount(row[keep], minlength=n)
    np.testing.assert_array_equal(got_n, want_n)
    np.testing.assert_array_equal(got, want)
    return int(ok.sum())


def test_recorded_lists_of_the_developed_dam_break(oracle):
    """The 401,856-particle dam break 30 frames in (the benchmark's state),
    WCSPH with its default launch shapes: every particle's recorded list
    against the oracle's within-h set."""
    from fluid_simulation_amd.examples.dam_break.scene import make_scene
    from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

    sim = DamBreakSimulation(make_scene(400000), "wcsph", device="cuda:0")
    for _ in range(30):
        sim.step()
    torch.cuda.synchronize()
    checked = _check_lists_vectorized(sim.solver, sim.model, sim.solver.smoothing_length, oracle)
    print(f"dam break: {checked} lists match", flush=True)
    assert checked > 0.99 * sim.model.particle_count
