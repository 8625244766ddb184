"""Pins of the oracle's APIC particle separation
(solver/apic/solver_apic.py:193-268, composed as :1003-1066).

Closed forms fix the push (strength * (rest - d) along the pair direction,
averaged over the pushing neighbours), the into-wall kill and the clamp; an
O(N^2) FP64 restatement that knows nothing of the grid or the visit order
pins the neighbour set and the Jacobi (all displacements, then all updates)
composition, including a grid with aliased buckets.
"""
import numpy as np
import pytest

from oracle import oracle as O

BOX = dict(bound_lo=(-1.0, -1.0, 0.0), bound_hi=(1.0, 1.0, 2.0))


def _params(rest=0.1, strength=0.55, iterations=1, wall_margin=0.01, wall_push_band=-1.0, radius=0.025, n=2,
            box=BOX):
    cfg = O.SeparationConfig(separation_rest_dist=rest, separation_strength=strength,
                             separation_iterations=iterations, wall_margin=wall_margin,
                             wall_push_band=wall_push_band, **box)
    return O.separation_params(cfg, np.full(n, radius, np.float32))


def _brute(x, flags, p):
    """Independent restatement: every pair, no grid, float64."""
    x = np.array(x, np.float64)
    lo, hi = np.array(p["bound_lo"], np.float64), np.array(p["bound_hi"], np.float64)
    rest, strength, band, margin = p["rest_dist"], p["strength"], p["wall_band"], p["wall_margin"]
    act = (flags & 1) != 0
    delta = np.zeros_like(x)
    for _ in range(p["iterations"]):
        delta = np.zeros_like(x)
        rij = x[:, None, :] - x[None, :, :]
        d = np.linalg.norm(rij, axis=2)
        pair = (d > 1e-8) & (d < rest) & act[:, None] & act[None, :]
        np.fill_diagonal(pair, False)
        with np.errstate(invalid="ignore", divide="ignore"):
            push = np.where(pair[:, :, None], rij / d[:, :, None] * (strength * (rest - d))[:, :, None], 0.0)
        cnt = pair.sum(axis=1)
        disp = np.where(cnt[:, None] > 0, push.sum(axis=1) / np.maximum(cnt, 1)[:, None], 0.0)
        for a in range(3):
            disp[(x[:, a] <= lo[a] + band) & (disp[:, a] < 0), a] = 0.0
            disp[(x[:, a] >= hi[a] - band) & (disp[:, a] > 0), a] = 0.0
        disp[~act] = 0.0
        delta = disp
        x = np.where(act[:, None], np.clip(x + disp, lo + margin, hi - margin), x)
    return x, delta


def test_two_particles_closed_form():
    p = _params()
    x = np.array([[0.0, 0.0, 1.0], [0.06, 0.0, 1.0]], np.float64)
    x1, d = O.particle_separation(x, np.ones(2, np.int32), p, precision="f64")
    push = 0.55 * (0.1 - 0.06)
    np.testing.assert_allclose(d, [[-push, 0, 0], [push, 0, 0]], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(x1, x + d, rtol=0, atol=1e-15)


def test_average_over_pushing_neighbours():
    """Middle of three collinear particles: pushes of 0.02 and 0.04 from both
    sides average to (0.55*0.04 - 0.55*0.02)/2 ... along +x from the nearer."""
    p = _params()
    x = np.array([[-0.08, 0, 1.0], [0.0, 0, 1.0], [0.06, 0, 1.0]], np.float64)
    _, d = O.particle_separation(x, np.ones(3, np.int32), p, precision="f64")
    # from the left (d=0.08): +x push 0.55*0.02; from the right (d=0.06): -x push 0.55*0.04
    np.testing.assert_allclose(d[1], [(0.55 * 0.02 - 0.55 * 0.04) / 2, 0, 0], atol=1e-15)
    # the outer pair (0.14 apart) does not interact
    np.testing.assert_allclose(d[0], [-0.55 * 0.02, 0, 0], atol=1e-15)


def test_into_wall_component_removed_and_clamp():
    p = _params(wall_margin=0.01)          # band = max(1.5m, 2m) = 0.02
    assert p["wall_band"] == pytest.approx(0.02)
    # particle 0 sits 0.015 above the floor and is pushed down by particle 1:
    # the -z component is removed, the +x one kept
    x = np.array([[0.0, 0.0, 0.015], [0.03, 0.0, 0.055]], np.float64)
    _, d = O.particle_separation(x, np.ones(2, np.int32), p, precision="f64")
    assert d[0, 2] == 0.0 and d[0, 0] < 0.0
    assert d[1, 2] > 0.0
    # a particle pushed past hi - margin is clamped there
    x = np.array([[0.985, 0.0, 1.0], [0.95, 0.0, 1.0]], np.float64)
    p2 = _params(wall_margin=0.01, wall_push_band=0.0)
    p2["wall_band"] = 0.0
    x1, d = O.particle_separation(x, np.ones(2, np.int32), p2, precision="f64")
    assert d[0, 0] > 0.0 and x1[0, 0] == pytest.approx(1.0 - 0.01)


def test_inactive_particles_neither_move_nor_push():
    p = _params()
    x = np.array([[0.0, 0.0, 1.0], [0.05, 0.0, 1.0], [0.0, 0.05, 1.0]], np.float64)
    flags = np.array([1, 0, 1], np.int32)
    x1, d = O.particle_separation(x, flags, p, precision="f64")
    np.testing.assert_array_equal(x1[1], x[1])
    np.testing.assert_array_equal(d[1], 0.0)
    assert d[0, 0] == 0.0 and d[0, 1] < 0.0        # pushed only by particle 2


@pytest.mark.parametrize("dims_override,iterations", [(None, 1), (None, 2), ((4, 4, 4), 2)])
def test_oracle_matches_brute_force(dims_override, iterations):
    rng = np.random.default_rng(7)
    n = 400
    x = rng.uniform([-0.3, -0.3, 0.0], [0.3, 0.3, 0.4], (n, 3))
    flags = (rng.random(n) > 0.1).astype(np.int32)
    p = _params(rest=0.06, iterations=iterations, n=n)
    if dims_override:
        p["dims"] = dims_override       # every bucket aliased many times over
    got, gd = O.particle_separation(x, flags, p, precision="f64")
    want, wd = _brute(x, flags, p)
    assert np.abs(wd).max() > 1e-3      # the case pushes
    np.testing.assert_allclose(gd, wd, rtol=0, atol=1e-13)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_default_parameters_follow_the_reference():
    cfg = O.SeparationConfig(bound_lo=(-1.2, -0.28, 0.0), bound_hi=(1.2, 0.28, 1.2))
    radius = np.full(100, 0.0039, np.float32)
    p = O.separation_params(cfg, radius)
    m = float(np.mean(radius))
    assert p["wall_margin"] == m                          # wall_margin < 0 -> mean radius
    assert p["wall_band"] == max(1.5 * m, 2.0 * m)
    assert p["rest_dist"] == 2.0 * m and p["query_radius"] == 2.0 * m
    assert p["iterations"] == 2 and p["strength"] == 0.55
    assert p["dims"] == tuple(int(v) for v in np.maximum(4, np.ceil(np.array([2.4, 0.56, 1.2]) / (2 * m))))
    assert O.separation_params(cfg, np.zeros(0, np.float32))["wall_margin"] == cfg.clamp_eps
