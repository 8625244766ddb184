"""The closed-form pins catch plausible slips (CPU only).

Each case compiles a copy of oracle/fs_oracle.c with one planted mistake (a
wrong constant, a dropped term, a swapped operand, the other float32
association) into a temporary library, runs the pins of
tests/test_oracle_closed_forms.py against it and requires at least one of
them to fail.  The committed oracle passes all of them."""
import os
import subprocess

import pytest

import tests.test_oracle_closed_forms as pins

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "fs_oracle.c")

VISC = ("R s = RC(10.0) * mu * (mass_nei / rmax(rho_nei, RC(1.0e-6))) * v_current_nei\n"
        "          / (distance * distance + RC(0.01) * (h * h));\n    a = vadd(a, smul(s, cubic_grad(d, h)));")
MUTATIONS = {
    "viscosity 10 -> 9": (VISC, VISC.replace("RC(10.0)", "RC(9.0)")),
    "viscosity 0.01 h^2 -> 0.001 h^2": (VISC, VISC.replace("RC(0.01)", "RC(0.001)")),
    "pressure p_j term dropped": ("return vadd(a, smul(-mass_nei * (dp_i + dp_nei), cubic_grad(d, h)));",
                                  "return vadd(a, smul(-mass_nei * (dp_i + dp_i), cubic_grad(d, h)));"),
    "pressure rho_j not squared": ("R dp_nei = p_nei / rmax(rho_nei * rho_nei, RC(1.0e-6));",
                                   "R dp_nei = p_nei / rmax(rho_nei * RC(1000.0), RC(1.0e-6));"),
    "EOS -1 -> -0.999": ("exponent[i]) - RC(1.0));", "exponent[i]) - RC(0.999));"),
    "factor sum_grad_sq halved": ("R denom = vdot(sum_grad, sum_grad) + sum_grad_sq;",
                                  "R denom = vdot(sum_grad, sum_grad) + sum_grad_sq * RC(0.5);"),
    "kappa dt^2 -> dt": ("kappa[t] = rmin((rho_err / (dt * dt)) * factor[t], max_kappa);",
                         "kappa[t] = rmin((rho_err / (dt)) * factor[t], max_kappa);"),
    "kappa_v rho dropped": ("(RC(1.0) / dt) * density_change * rho[t] * factor[t]",
                            "(RC(1.0) / dt) * density_change * factor[t]"),
    "correction kappa_j -> kappa_i": ("kappa[i] / rmax(rho[i], RC(1.0e-6)) + kappa[j] / rmax(rho[j], RC(1.0e-6))",
                                      "kappa[i] / rmax(rho[i], RC(1.0e-6)) + kappa[i] / rmax(rho[j], RC(1.0e-6))"),
    "lambda |sum grad|^2 dropped": ("        sum_grad_sq += vdot(grad_i, grad_i);\n", "\n"),
    "lambda epsilon halved": ("-c_i / (sum_grad_sq + lambda_epsilon)", "-c_i / (sum_grad_sq + RC(0.5) * lambda_epsilon)"),
    "delta p lambda_j -> lambda_i": ("(pbf_lambda[i] + pbf_lambda[j] + scorr)", "(pbf_lambda[i] + pbf_lambda[i] + scorr)"),
    "scorr exponent n -> n-1": ("return -scorr_k * rpow(ratio, scorr_n);", "return -scorr_k * rpow(ratio, scorr_n - RC(1.0));"),
    "XSPH c halved": ("vstore(v_out, i, vadd(vi, smul(xsph_c, corr)));",
                      "vstore(v_out, i, vadd(vi, smul(xsph_c * RC(0.5), corr)));"),
    "cohesion (g/m*m_j)*(d*W)": ("a = vsub(a, muls(smul(gamma / mass * mass_nei, d), cubic_w(e_dist, h)));",
                                 "a = vsub(a, smul(gamma / mass * mass_nei, muls(d, cubic_w(e_dist, h))));"),
}

PINS = [
    lambda o: [pins.test_viscosity_term_is_the_sph_laplacian(f, o) for f in ("div_free", "xx", "xy")],
    pins.test_pressure_force_is_minus_grad_p_over_rho,
    lambda o: [pins.test_tait_eos_on_the_lattice_density(r, c, o) for r, c in ((1.02, True), (0.97, False))],
    pins.test_cohesion_association_is_warps,
    pins.test_dfsph_factor_three_particles,
    pins.test_dfsph_density_solve_two_particles,
    pins.test_dfsph_divergence_solve_two_particles,
    lambda o: [pins.test_pbf_iteration_two_particles(v, o) for v in ("plain", "scorr", "clamp")],
    pins.test_pbf_xsph_of_a_linear_velocity_field_vanishes_inside,
]


def _build(src_text, tmp_path, precision):
    c = tmp_path / "fs_oracle_mut.c"
    c.write_text(src_text)
    so = tmp_path / f"libmut_{precision}.so"
    real = "float" if precision == "f32" else "double"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-std=c11",
                    f"-DORACLE_REAL={real}", "-o", str(so), str(c), "-lm"], check=True, capture_output=True)
    return str(so)


def _failures(oracle):
    n = 0
    for pin in PINS:
        try:
            pin(oracle)
        except AssertionError:
            n += 1
    return n


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_a_planted_mistake_fails_a_pin(name, oracle, tmp_path, monkeypatch):
    old, new = MUTATIONS[name]
    src = open(SRC).read()
    assert src.count(old) == 1, "mutation site not found"
    mutated = src.replace(old, new)
    libs = {p: oracle.load_library(_build(mutated, tmp_path, p), p) for p in ("f32", "f64")}
    monkeypatch.setattr(oracle, "_LIBS", libs)
    assert _failures(oracle) >= 1, f"no pin fails with the planted mistake: {name}"
