"""Closed-form pins for the oracle's force, pressure and solver formulas (CPU).

Each test fixes an oracle function by what the mathematics gives, not by
re-evaluating the oracle's own expression:

* the viscosity term of cal_acc_with_non_pressure (solver_wcsph.py:79-88) is
  Monaghan's SPH Laplacian: for a smooth field its sum tends to
  -(J/3) mu (lap v + 2 grad div v), J = int r^3 W'(r) / (r^2 + 0.01 h^2) dV,
  which a lattice with a wide support reproduces to 0.1 %;
* the symmetric pressure force (solver_wcsph.py:92-106) tends to -grad p / rho
  on a linear pressure field;
* the Tait EOS (solver_wcsph.py:151-156) on the lattice density of the
  closed-form lattice sum;
* DFSPH's factor, kappa, kappa_v, density-change rate and both velocity
  corrections (solver_dfsph.py:136-174, 245-403), and PBF's lambda, delta p,
  scorr, clamp and XSPH (solver_pbf.py:90-101, 245-463), on two- and
  three-particle configurations written out by hand with the textbook kernel
  derivatives;
* the float32 association of the cohesion term, ((gamma/m_i*m_j)*d)*W
  (solver_wcsph.py:78), which Warp evaluates left to right.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate

# --------------------------------------------------------------------------- textbook kernels (float64)


def cubic_w(r, h):
    """Monaghan cubic spline with support h (k = 8 / (pi h^3))."""
    k = 8.0 / (math.pi * h ** 3)
    q = r / h
    if q <= 0.5:
        return k * (6 * q ** 3 - 6 * q ** 2 + 1)
    return 2 * k * (1 - q) ** 3 if q <= 1 else 0.0


def cubic_dw(r, h):
    """dW/dr of cubic_w, differentiated branch by branch."""
    k = 8.0 / (math.pi * h ** 3)
    q = r / h
    if q <= 0.5:
        return k * (18 * q * q - 12 * q) / h
    return -6 * k * (1 - q) ** 2 / h if q <= 1 else 0.0


def cubic_grad(rv, h):
    r = float(np.linalg.norm(rv))
    return np.zeros(3) if r == 0.0 else cubic_dw(r, h) * np.asarray(rv) / r


def poly6(r, h):
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3 if r < h else 0.0


def spiky_grad(rv, h):
    """Gradient of the spiky kernel 15/(pi h^6) (h - r)^3 (Mueller et al. 2003)."""
    r = float(np.linalg.norm(rv))
    if r == 0.0 or r >= h:
        return np.zeros(3)
    return -45.0 / (math.pi * h ** 6) * (h - r) ** 2 * np.asarray(rv) / r


def _grid(oracle, x, h, precision="f64"):
    return oracle.OracleGrid(precision=precision).build(x, h)


# --------------------------------------------------------------------------- WCSPH forces and EOS


def _wide_lattice(coff, d=0.01):
    half = int(2 * coff) + 4
    ax = np.arange(-half, half + 1) * d
    pts = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) + np.array([0.0, 0.0, 0.5])
    c = pts - np.array([0.0, 0.0, 0.5])
    centre = np.all(np.abs(c) < 1.5 * d, axis=1)
    return pts, c, centre


@pytest.mark.parametrize("field", ["div_free", "xx", "xy"])
def test_viscosity_term_is_the_sph_laplacian(field, oracle):
    """Interior particles of a lattice with h = 6 spacings: the viscosity
    acceleration of a quadratic velocity field equals the continuum limit of
    the term, -(J/3) mu (lap v + 2 grad div v).  A wrong 10, a wrong
    0.01 h^2 (0.001 h^2 moves the limit by 4 %) or a dropped m_j/rho_j fails."""
    d, coff, mu = 0.01, 6.0, 1.0e-3
    h = coff * d
    pts, c, centre = _wide_lattice(coff, d)
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    zero = 0.0 * x
    v, lap_plus = {"div_free": (np.stack([y * y, z * z, x * x], 1), [2.0, 2.0, 2.0]),   # div v = 0
                   "xx": (np.stack([x * x, zero, zero], 1), [2.0 + 4.0, 0.0, 0.0]),     # lap + 2 grad div
                   "xy": (np.stack([x * y, zero, zero], 1), [0.0, 2.0, 0.0])}[field]
    m = oracle.uniform_model(len(pts), 1000.0 * d ** 3, stiffness=0.0, exponent=7.0, viscosity=mu,
                             gravity=(0, 0, 0), dtype=np.float64)
    cfg = oracle.WcsphConfig(particle_radius=0.5 * d, particle_length=d, smoothing_length_coff=coff,
                             boundary_damping=0.0, bound_width=5, bound_height=5, bound_length=5)
    _, ex = oracle.wcsph_step(m, oracle.OState(pts, v), cfg, 1e-6, "f64")
    eps = 0.01 * h * h
    J, _ = integrate.quad(lambda r: 4 * math.pi * r ** 5 * cubic_dw(r, h) / (r * r + eps), 0.0, h,
                          points=[0.5 * h], epsabs=1e-15)
    expect = -(10.0 / 30.0) * J * mu * np.asarray(lap_plus)
    got = ex["accel"][centre]
    scale = np.abs(expect).max()
    np.testing.assert_allclose(got, np.broadcast_to(expect, got.shape), rtol=0, atol=3e-3 * scale)


def test_pressure_force_is_minus_grad_p_over_rho(oracle):
    """The symmetric pressure term on a linear pressure field: interior
    particles of a wide lattice accelerate by -grad p / rho (to 0.3 %).  A
    p/rho instead of p/rho^2, a dropped p_j term or a sign slip fails."""
    d, coff = 0.01, 6.0
    h = coff * d
    pts, c, centre = _wide_lattice(coff, d)
    n = len(pts)
    lib = oracle._lib("f64")
    R = lib._real
    grad_p = np.array([5000.0, 10000.0, -5000.0])
    pressure = np.ascontiguousarray(20000.0 + c @ grad_p)
    m = oracle.uniform_model(n, 500.0 * d ** 3, rest_density=500.0, gravity=(0, 0, 0), dtype=np.float64)
    g = _grid(oracle, pts, h)
    gs = g.struct()
    vol = np.full(n, d ** 3)                             # rest_density * volume = mass
    rho = np.zeros(n)
    lib.ora_density(ctypes.byref(gs), oracle._p(m.flags), oracle._p(pts), oracle._p(m.rest_density),
                    oracle._p(vol), R(h), ctypes.c_int(0), oracle._p(rho))
    accel = np.zeros((n, 3))
    zeros = np.zeros(n)
    lib.ora_wcsph_acceleration(ctypes.byref(gs), oracle._p(m.flags), oracle._p(m.mask), oracle._p(pts),
                               oracle._p(np.zeros((n, 3))), oracle._p(rho), oracle._p(pressure), R(0.5 * d),
                               oracle._p(m.mass), oracle._p(zeros), oracle._p(zeros), oracle._p(m.gravity),
                               oracle._p(m.world), R(h), oracle._p(accel))
    expect = -grad_p[None, :] / rho[centre][:, None]
    np.testing.assert_allclose(accel[centre], expect, rtol=0, atol=3e-3 * np.abs(expect).max())


@pytest.mark.parametrize("ratio, clamp", [(1.02, True), (0.97, True), (0.97, False)])
def test_tait_eos_on_the_lattice_density(ratio, clamp, oracle):
    """p = B((rho/rho0)^gamma - 1), clamped at 0 (solver_wcsph.py:151-156),
    with rho the closed-form lattice sum and rho0 = rho/ratio."""
    d = 0.01
    h = 2 * d
    n_side = 9
    ax = np.arange(n_side) * d
    pts = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) + np.array([0.003, 0.003, 0.203])
    mass = 1000.0 * d ** 3
    o = np.arange(-2, 3)
    offs = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3) * d
    rho_c = mass * sum(cubic_w(np.linalg.norm(v), h) for v in offs)
    rho0 = rho_c / ratio
    B, gam = 250000.0, 7.0
    m = oracle.uniform_model(len(pts), mass, rest_density=rho0, stiffness=B, exponent=gam, gravity=(0, 0, 0),
                             dtype=np.float64)
    cfg = oracle.WcsphConfig(particle_radius=0.5 * d, particle_length=d, smoothing_length_coff=2.0,
                             boundary_damping=0.0, bound_width=1.2, bound_height=0.28, bound_length=0.6,
                             clamp_negative_pressure=clamp)
    out, _ = oracle.wcsph_step(m, oracle.OState(pts, np.zeros_like(pts)), cfg, 1e-6, "f64")
    centre = (n_side ** 3) // 2
    expect = B * (ratio ** gam - 1.0)
    if clamp:
        expect = max(expect, 0.0)
    assert out.sph["pressure"][centre] == pytest.approx(expect, rel=1e-9, abs=1e-9)
    if ratio > 1:
        assert expect == pytest.approx(B * 0.14868566764928, rel=1e-12)     # 1.02^7 - 1 = 0.14868566764928


def test_cohesion_association_is_warps(oracle):
    """Isolated pairs, no pressure and no viscosity: the float32 acceleration
    is exactly 0 - ((gamma/m_i*m_j)*d)*W(max(r, radius)) evaluated left to
    right, as Warp evaluates `gamma / mass * mass_nei * d * get_cubic(...)`.
    The other association, (gamma/m_i*m_j)*(d*W), differs for some pairs."""
    f32 = np.float32
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
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32) * f32(1000.0 * d ** 3)
    gamma = rng.uniform(0.001, 0.05, n).astype(np.float32)
    m = oracle.uniform_model(n, 1.0, stiffness=0.0, exponent=7.0, viscosity=0.0, gravity=(0, 0, 0))
    m.mass[:] = mass
    m.surface_tension[:] = gamma
    cfg = oracle.WcsphConfig(particle_radius=0.5 * d, particle_length=d, smoothing_length_coff=2.0,
                             boundary_damping=0.0, bound_width=5, bound_height=5, bound_length=5)
    _, ex = oracle.wcsph_step(m, oracle.OState(pts, np.zeros_like(pts)), cfg, 1e-6)
    lib = oracle._lib("f32")
    R = lib._real
    radius = f32(0.5 * d)
    hf = f32(h)
    partner = np.concatenate([np.arange(npairs) + npairs, np.arange(npairs)])
    warp = np.zeros((n, 3), np.float32)
    other = np.zeros((n, 3), np.float32)
    for i in range(n):
        j = partner[i]
        dv = pts[i] - pts[j]
        dist = np.sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
        w = f32(lib.ora_cubic_w(R(float(max(dist, radius))), R(float(hf))))
        s = (gamma[i] / mass[i]) * mass[j]
        warp[i] = f32(0.0) - (s * dv) * w
        other[i] = f32(0.0) - s * (dv * w)
    np.testing.assert_array_equal(ex["accel"], warp)
    assert np.any(warp != other)


# --------------------------------------------------------------------------- DFSPH


def _dfsph_cfg(oracle, d, **kw):
    base = dict(max_rho_iterations=1, max_vel_iterations=0, warm_start=False, max_kappa=1.0e30,
                max_kappa_v=1.0e30, min_neighbors_for_factor=1, particle_radius=0.5 * d, particle_length=d,
                smoothing_length_coff=2.0, boundary_damping=0.0, bound_width=5, bound_height=5, bound_length=5)
    base.update(kw)
    return oracle.DfsphConfig(**base)


def test_dfsph_factor_three_particles(oracle):
    """factor_i = 1 / (|sum_j m_j grad W_ij|^2 + sum_j |m_j grad W_ij|^2)
    (solver_dfsph.py:150-166) for three particles at rest in an under-dense
    state (kappa = 0, so positions do not move)."""
    d = 0.01
    h = 2 * d
    x = np.array([[0.0, 0.0, 0.3], [0.7 * h, 0.0, 0.3], [0.1 * h, 0.35 * h, 0.3 + 0.2 * h]])
    mass = np.array([1.0e-3, 1.3e-3, 0.8e-3])
    m = oracle.uniform_model(3, 1.0, rest_density=1.0e9, gravity=(0, 0, 0), viscosity=0.0, dtype=np.float64)
    m.mass[:] = mass
    out, _ = oracle.dfsph_step(m, oracle.OState(x, np.zeros_like(x)), _dfsph_cfg(oracle, d), 1e-3, "f64")
    for i in range(3):
        grads = [mass[j] * cubic_grad(x[i] - x[j], h) for j in range(3) if j != i]
        s = np.sum(grads, axis=0)
        expect = 1.0 / (s @ s + sum(g @ g for g in grads))
        assert out.sph["factor"][i] == pytest.approx(expect, rel=1e-12)
        assert out.sph["neighbor_count"][i] == 2
    np.testing.assert_array_equal(out.x, x)


def _approaching_pair(h, u=0.3):
    x = np.array([[-0.3 * h, 0.0, 0.3], [0.3 * h, 0.02 * h, 0.3]])
    v = np.array([[u, 0.0, 0.0], [-u, 0.05, 0.0]])
    return x, v


def _pair_by_hand(x, v, mass, h):
    """rho, factor, grad W_ij (per particle) and the rate D rho of a pair."""
    g01 = cubic_grad(x[0] - x[1], h)
    g = [g01, -g01]
    r = np.linalg.norm(x[0] - x[1])
    other = (1, 0)
    rho = np.array([mass[i] * cubic_w(0.0, h) + mass[other[i]] * cubic_w(r, h) for i in (0, 1)])
    factor = np.array([1.0 / (2.0 * mass[other[i]] ** 2 * (g01 @ g01)) for i in (0, 1)])
    rate = np.array([mass[other[i]] * (v[i] - v[other[i]]) @ g[i] for i in (0, 1)])
    return rho, factor, g, rate


def _pair_correction(v, mass, kap, rho, g, dt):
    out = v.copy()
    for i, j in ((0, 1), (1, 0)):
        out[i] = v[i] - dt * mass[j] * (kap[i] / rho[i] + kap[j] / rho[j]) * g[i]
    return out


def test_dfsph_density_solve_two_particles(oracle):
    """One constant-density iteration by hand for two particles of unequal
    mass: rho, the rate D rho_i = sum_j m_j (v_i - v_j).grad W_ij (:245-272),
    rho* = rho + dt D rho (:275-287), kappa = max(rho* - rho0, 0)/dt^2 * factor
    (:290-306), the correction v*_i -= dt sum_j m_j (kappa_i/rho_i +
    kappa_j/rho_j) grad W_ij (:309-337) and x' = x + dt v* (:340-353)."""
    d = 0.01
    h = 2 * d
    dt = 1.0e-3
    mass = np.array([1.0e-3, 1.6e-3])
    rho0 = 1.0
    x, v = _approaching_pair(h)
    m = oracle.uniform_model(2, 1.0, rest_density=rho0, gravity=(0, 0, 0), viscosity=0.0, dtype=np.float64)
    m.mass[:] = mass
    out, ex = oracle.dfsph_step(m, oracle.OState(x, v), _dfsph_cfg(oracle, d), dt, "f64")
    rho, factor, g, rate = _pair_by_hand(x, v, mass, h)
    rho_star = rho + dt * rate
    kappa = np.maximum(rho_star - rho0, 0.0) / dt ** 2 * factor
    v_star = _pair_correction(v, mass, kappa, rho, g, dt)
    np.testing.assert_allclose(out.sph["d_rho_dt"], rate, rtol=1e-12)
    np.testing.assert_allclose(out.sph["rho_star"], rho_star, rtol=1e-12)
    np.testing.assert_allclose(out.sph["kappa"], kappa, rtol=1e-12)
    np.testing.assert_allclose(out.v, v_star, rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(out.x, x + dt * v_star, rtol=1e-12)
    assert ex["rho_iterations"] == 1 and ex.get("rho_corrections") == 1
    assert kappa.min() > 0.0 and abs(kappa[0] - kappa[1]) > 1e-3 * kappa.max()


def test_dfsph_divergence_solve_two_particles(oracle):
    """One divergence-free iteration by hand at x' = x + dt v (:769-883):
    kappa_v = (1/dt) max(D rho', 0) rho' factor' (:356-372) and the
    correction of :375-403, with rho' and factor' at x'."""
    d = 0.01
    h = 2 * d
    dt = 1.0e-3
    mass = np.array([1.0e-3, 1.6e-3])
    x, v = _approaching_pair(h)
    m = oracle.uniform_model(2, 1.0, rest_density=1.0, gravity=(0, 0, 0), viscosity=0.0, dtype=np.float64)
    m.mass[:] = mass
    cfg = _dfsph_cfg(oracle, d, max_rho_iterations=0, max_vel_iterations=1)
    out, ex = oracle.dfsph_step(m, oracle.OState(x, v), cfg, dt, "f64")
    xp = x + dt * v
    rho, factor, g, rate = _pair_by_hand(xp, v, mass, h)
    kappa_v = (1.0 / dt) * np.maximum(rate, 0.0) * rho * factor
    v_out = _pair_correction(v, mass, kappa_v, rho, g, dt)
    np.testing.assert_allclose(out.x, xp, rtol=1e-12)
    np.testing.assert_allclose(out.sph["rho"], rho, rtol=1e-12)
    np.testing.assert_allclose(out.sph["factor"], factor, rtol=1e-12)
    np.testing.assert_allclose(out.sph["d_rho_dt"], rate, rtol=1e-12)
    np.testing.assert_allclose(out.sph["kappa_v"], kappa_v, rtol=1e-12)
    np.testing.assert_allclose(out.v, v_out, rtol=1e-11, atol=1e-14)
    assert kappa_v.min() > 0.0 and abs(kappa_v[0] - kappa_v[1]) > 1e-3 * kappa_v.max()
    assert ex["vel_iterations"] == 1


# --------------------------------------------------------------------------- PBF


def _pbf_cfg(oracle, d, **kw):
    base = dict(max_iterations=1, lambda_epsilon=100.0, min_neighbors_for_lambda=0, scorr_k=0.0,
                max_delta_position=0.0, xsph_c=0.05, particle_radius=0.5 * d, particle_length=d,
                smoothing_length_coff=2.0, boundary_damping=0.0, bound_width=5, bound_height=5, bound_length=5)
    base.update(kw)
    return oracle.PbfConfig(**base)


@pytest.mark.parametrize("variant", ["plain", "scorr", "clamp"])
def test_pbf_iteration_two_particles(variant, oracle):
    """One PBF iteration on a compressed pair of unequal masses at rest, by
    hand: rho_i = sum_j m_j poly6, C = rho/rho0 - 1, grad_j = -(m_j/rho0) grad
    spiky, lambda = -C / (sum |grad_j|^2 + |sum grad_j|^2 + eps) (:245-302);
    dp_i = (1/rho0) sum_{j!=i} (lambda_i + lambda_j + scorr) m_j grad spiky
    (:305-350), scorr = -k (poly6(r)/poly6(q h))^n (:90-101), |dp| clamped
    (:344-348); v = dp/dt (:402-416); XSPH v_i += c sum (m_j/rho_j)(v_j - v_i)
    poly6 at the final positions (:434-463)."""
    d = 0.01
    h = 2 * d
    dt = 1.0e-3
    # eps keeps |dp| below h/4, so both particles stay in the query range of
    # the grid built before the iteration (which the final density uses)
    rho0, eps, c = 1.0, 2.0e10, 0.05
    mass = np.array([1.0e-3, 1.5e-3])
    x = np.array([[0.0, 0.0, 0.3], [0.45 * h, 0.1 * h, 0.3]])
    kw = {"plain": {}, "scorr": dict(scorr_k=1.0e-7, scorr_n=4.0, scorr_q=0.3),
          "clamp": dict(max_delta_position=1.0e-6)}[variant]
    m = oracle.uniform_model(2, 1.0, rest_density=rho0, gravity=(0, 0, 0), viscosity=0.0, dtype=np.float64)
    m.mass[:] = mass
    out, _ = oracle.pbf_step(m, oracle.OState(x, np.zeros_like(x)),
                             _pbf_cfg(oracle, d, lambda_epsilon=eps, **kw), dt, "f64")
    other = (1, 0)
    r = np.linalg.norm(x[0] - x[1])
    gs = [spiky_grad(x[0] - x[1], h), spiky_grad(x[1] - x[0], h)]
    lam = np.zeros(2)
    for i in (0, 1):
        j = other[i]
        rho = mass[i] * poly6(0.0, h) + mass[j] * poly6(r, h)
        grad_j = -(mass[j] / rho0) * gs[i]
        lam[i] = -(rho / rho0 - 1.0) / (2.0 * (grad_j @ grad_j) + eps)
    scorr = 0.0
    if variant == "scorr":
        scorr = -1.0e-7 * (poly6(r, h) / poly6(0.3 * h, h)) ** 4.0
    dp = np.array([(1.0 / rho0) * (lam[0] + lam[1] + scorr) * mass[other[i]] * gs[i] for i in (0, 1)])
    if variant == "clamp":
        dp = dp * (1.0e-6 / np.linalg.norm(dp, axis=1, keepdims=True))
    xo = x + dp
    ro = np.linalg.norm(xo[0] - xo[1])
    rho_f = np.array([mass[i] * poly6(0.0, h) + mass[other[i]] * poly6(ro, h) for i in (0, 1)])
    vel = dp / dt
    v_out = np.array([vel[i] + c * (mass[other[i]] / rho_f[other[i]]) * (vel[other[i]] - vel[i]) * poly6(ro, h)
                      for i in (0, 1)])
    np.testing.assert_allclose(out.sph["pbf_lambda"], lam, rtol=1e-12)
    np.testing.assert_allclose(out.sph["delta_p"], dp, rtol=1e-11)
    np.testing.assert_allclose(out.x, xo, rtol=1e-12)
    np.testing.assert_allclose(out.sph["rho"], rho_f, rtol=1e-12)
    np.testing.assert_allclose(out.v, v_out, rtol=1e-6)
    np.testing.assert_array_equal(out.sph["neighbor_count"], [1, 1])
    assert lam.max() < 0.0 and abs(lam[0] - lam[1]) > 1e-3 * abs(lam).max()
    assert 0.0 < np.abs(dp).max() < 0.25 * h
    if variant == "scorr":
        assert abs(scorr) > 0.1 * abs(lam).max()


def test_pbf_xsph_of_a_linear_velocity_field_vanishes_inside(oracle):
    """Under-dense lattice (lambda = 0, no position correction) moving with
    a linear velocity field: XSPH adds c sum (m_j/rho_j)(v_j - v_i) W_ij, whose
    first moment vanishes on the symmetric interior stencil."""
    d = 0.02
    ax = np.arange(9) * (1.2 * d)                       # a float64 lattice: exactly symmetric stencils
    pts = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) + np.array([0.0, 0.0, 0.3])
    n = len(pts)
    A = np.array([[0.0, 0.3, -0.1], [0.2, 0.1, 0.0], [-0.3, 0.0, 0.2]])
    v = (pts - pts.mean(0)) @ A.T
    m = oracle.uniform_model(n, 1000 * d ** 3, viscosity=0.0, gravity=(0, 0, 0), dtype=np.float64)
    dt = 1.0e-4
    out, _ = oracle.pbf_step(m, oracle.OState(pts, v), _pbf_cfg(oracle, d, xsph_c=0.5), dt, "f64")
    np.testing.assert_array_equal(out.sph["pbf_lambda"], 0.0)
    c = pts - pts.mean(0)
    inner = np.all(np.abs(c) < 2.5 * 1.2 * d, axis=1)
    v_pos = (pts + dt * v - pts) / dt
    np.testing.assert_allclose(out.v[inner], v_pos[inner], rtol=0, atol=1e-9)
    # ... while at the boundary the one-sided stencil does pull the velocities
    assert np.abs(out.v[~inner] - v_pos[~inner]).max() > 1e-4
