"""Pins for the oracle's WCSPH adjoint (CPU only): the reverse sweep of a short
FP64 rollout agrees with FP64 central finite differences of the forward
oracle for every parameter the adjoint returns: stiffness, viscosity, initial
velocities and initial positions.  The scenes are compressed blocks with
random velocities, so pressure, viscosity and wall clamps all carry gradient.
A dropped or sign-flipped term in any of the three adjoint passes moves the
adjoint far outside the tolerance."""
from dataclasses import replace

import numpy as np
import pytest


def _block(n_side, d, squeeze, seed, z0=0.1):
    rng = np.random.default_rng(seed)
    g = np.arange(n_side) * d
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + np.array([0.0, 0.0, z0])
    x = (x + rng.uniform(-0.15 * d, 0.15 * d, x.shape)) * squeeze
    v = rng.normal(size=x.shape) * 0.3
    return x, v, rng


def _case(oracle, walls):
    d = 0.02
    x, v, rng = _block(6, d, 0.93, seed=1 if not walls else 2, z0=0.0015 if walls else 0.1)
    n = len(x)
    m = oracle.uniform_model(n, 1000 * d ** 3, stiffness=2.0e4, exponent=7.0, viscosity=5e-3, dtype=np.float64)
    if walls:
        # floor at z = 0 and a wall at x = -0.0015: the bottom layer and the
        # -x face get clamped, outward velocity is damped
        x[:, 0] -= x[:, 0].min() - 0.001
        v[:, 2] -= 0.8
        v[:, 0] -= 0.5
        cfg = oracle.WcsphConfig(particle_radius=d / 2, particle_length=d, smoothing_length_coff=2.0,
                                 bound_width=0.0015 + 10.0, bound_height=5.0, bound_length=5.0,
                                 boundary_damping=-0.3)
        x[:, 0] -= 10.0      # the -x wall at -bound_width sits just below the block
    else:
        cfg = oracle.WcsphConfig(particle_radius=d / 2, particle_length=d, smoothing_length_coff=2.0,
                                 bound_width=5.0, bound_height=5.0, bound_length=5.0, boundary_damping=0.0)
    wx, wv = rng.normal(size=x.shape), rng.normal(size=x.shape)
    return m, cfg, x, v, wx, wv, rng


def _loss(oracle, m, cfg, x0, v0, wx, wv, dt, steps, precision="f64"):
    states = oracle.wcsph_rollout(m, x0, v0, cfg, dt, steps, precision)
    xT, vT = states[-1]
    return float((wx * xT).sum() + (wv * vT).sum()), states


@pytest.mark.parametrize("walls", [False, True])
def test_wcsph_adjoint_matches_fp64_central_differences(oracle, walls):
    m, cfg, x, v, wx, wv, rng = _case(oracle, walls)
    dt, steps = 2e-4, 6
    _, states = _loss(oracle, m, cfg, x, v, wx, wv, dt, steps)
    xb, vb, gs, gv = oracle.wcsph_rollout_adjoint(m, states, cfg, dt, wx, wv, "f64")
    if walls:
        clamped = sum(int((s[0][:, 2] == 0.0).sum()) for s in states[1:])
        assert clamped > 0, "the wall case must exercise the clamp"
    n = m.n

    def fd(f, eps):
        return (f(eps) - f(-eps)) / (2.0 * eps)

    # steps sized so that the loss's rounding (positions reach |x| ~ 10 m in
    # the wall case) stays far below the difference; central differences
    # carry O(eps^2) truncation error
    B, nu = float(m.stiffness[0]), float(m.viscosity[0])
    d_stiff = fd(lambda e: _loss(oracle, replace(m, stiffness=np.full(n, B * (1 + e))), cfg, x, v, wx, wv, dt,
                                 steps)[0], 1e-5) / B
    d_visc = fd(lambda e: _loss(oracle, replace(m, viscosity=np.full(n, nu * (1 + e))), cfg, x, v, wx, wv, dt,
                                steps)[0], 1e-4) / nu
    dv = rng.normal(size=x.shape)
    d_v = fd(lambda e: _loss(oracle, m, cfg, x, v + e * dv, wx, wv, dt, steps)[0], 1e-4)
    dx = rng.normal(size=x.shape)
    d_x = fd(lambda e: _loss(oracle, m, cfg, x + e * dx, v, wx, wv, dt, steps)[0], 1e-6)
    assert gs.sum() == pytest.approx(d_stiff, rel=1e-6)
    assert gv.sum() == pytest.approx(d_visc, rel=1e-6)
    assert (vb * dv).sum() == pytest.approx(d_v, rel=1e-6)
    assert (xb * dx).sum() == pytest.approx(d_x, rel=1e-6)
    assert abs(gs.sum()) > 0 and abs(gv.sum()) > 0


def test_fp32_adjoint_tracks_fp64(oracle):
    m, cfg, x, v, wx, wv, _ = _case(oracle, False)
    dt, steps = 2e-4, 3
    _, s64 = _loss(oracle, m, cfg, x, v, wx, wv, dt, steps, "f64")
    _, s32 = _loss(oracle, m, cfg, x.astype(np.float32), v.astype(np.float32), wx, wv, dt, steps, "f32")
    a64 = oracle.wcsph_rollout_adjoint(m, s64, cfg, dt, wx, wv, "f64")
    a32 = oracle.wcsph_rollout_adjoint(m, s32, cfg, dt, wx, wv, "f32")
    for g64, g32 in zip(a64, a32):
        scale = np.abs(g64).max()
        assert np.abs(np.asarray(g32, np.float64) - g64).max() <= 2e-4 * scale
