"""Host-side checks (CPU): scene generator known answers, the Newton-compatible
shim, the solver configs and attributes against the reference's, and the C-ABI
library (loads, exports every declared symbol, struct layouts agree with the
header).  No GPU compute here."""
import ctypes
import os
import re
import subprocess
import sys
from dataclasses import asdict, fields

import numpy as np
import pytest

from fluid_simulation_amd import _lib
from fluid_simulation_amd.examples.dam_break.scene import DEFAULT_SPEC, make_scene
from fluid_simulation_amd.newton_compat import AttributeAssignment, Axis, Model, ModelBuilder, ParticleFlags
from fluid_simulation_amd.solver.pbf import SolverPBF
from fluid_simulation_amd.solver.sph.solver_dfsph import SolverDFSPH
from fluid_simulation_amd.solver.sph.solver_wcsph import SolverWCSPH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluidsim.h")


# ---------------------------------------------------------------- scene (unit_test_dam_break.py:19-27)

def test_default_scene_specifications():
    scene = make_scene()
    assert scene.count == 401856
    assert scene.report()["lattice_resolution"] == [78, 56, 92]
    np.testing.assert_allclose(scene.positions.min(axis=0), [-1.17, -0.25, 0.03], atol=1e-6)
    np.testing.assert_allclose(scene.positions.max(axis=0), [-0.468, 0.25, 0.8508], atol=1e-6)
    assert scene.radius * 2.0 == scene.spacing
    assert scene.mass == scene.spec.density * scene.volume
    assert scene.ideal_height == pytest.approx(scene.count * scene.mass / 1000.0 / 1.344)
    assert scene.ideal_height == pytest.approx(0.2194114502620966, rel=1e-12)   # performance_summary.json:29


def test_small_scene_counts():
    assert make_scene(512).count == 540
    assert make_scene(10000).count == 9936
    a, b = make_scene(512).report(), make_scene(512).report()
    assert a["initial_position_sha256"] == b["initial_position_sha256"]


# ---------------------------------------------------------------- configs and attributes

REF_WCSPH = dict(particle_radius=0.1, particle_length=0.2, smoothing_length_coff=1.35, bound_width=10.0,
                 bound_height=10.0, bound_length=10.0, boundary_damping=-0.1,
                 clamp_negative_pressure=True)                                   # solver_wcsph.py:309-318
REF_DFSPH = dict(max_rho_iterations=2, max_vel_iterations=1, warm_start=True, eta_density=0.1,
                 eta_divergence=1e-4, max_kappa=1e4, max_kappa_v=1e6, min_neighbors_for_factor=15,
                 particle_radius=0.1, particle_length=0.2, smoothing_length_coff=1.35, bound_width=10.0,
                 bound_height=10.0, bound_length=10.0, boundary_damping=-0.1, voxel_size=0.3)  # solver_dfsph.py:446-463
REF_PBF = dict(max_iterations=10, lambda_epsilon=100.0, min_neighbors_for_lambda=0, scorr_k=1e-4, scorr_n=4.0,
               scorr_q=0.3, max_delta_position=0.0, xsph_c=0.01, particle_radius=0.1, particle_length=0.2,
               smoothing_length_coff=1.35, bound_width=10.0, bound_height=10.0, bound_length=10.0,
               boundary_damping=0.0, boundary_epsilon=0.0, rebuild_grid_each_iteration=False)  # solver_pbf.py:521-539


@pytest.mark.parametrize("cfg,ref", [(SolverWCSPH.SphConfig, REF_WCSPH), (SolverDFSPH.SphConfig, REF_DFSPH),
                                     (SolverPBF.PbfConfig, REF_PBF)])
def test_config_fields_and_defaults_match_reference(cfg, ref):
    assert [f.name for f in fields(cfg)] == list(ref)
    assert asdict(cfg()) == ref
    # the benchmark's solver swap constructs Config(**asdict(config)) (benchmark.py:48)
    assert cfg(**asdict(cfg())) == cfg()


REF_ATTRS = {
    SolverWCSPH: ({"particle_mask": ("int32", 0), "rest_density": ("float32", 1000.0),
                   "surface_tension": ("float32", 0.01), "viscosity": ("float32", 0.005),
                   "stiffness": ("float32", 85000.0), "exponent": ("float32", 7.0)},
                  {"rho": ("float32", 0.0), "pressure": ("float32", 0.0)}),
    SolverDFSPH: ({"particle_mask": ("int32", 0), "rest_density": ("float32", 1000.0),
                   "surface_tension": ("float32", 0.01), "viscosity": ("float32", 0.005)},
                  {"rho": ("float32", 0.0), "factor": ("float32", 0.0), "rho_star": ("float32", 0.0),
                   "d_rho_dt": ("float32", 0.0), "kappa": ("float32", 0.0), "kappa_v": ("float32", 0.0),
                   "neighbor_count": ("int32", 0), "v_star": ("vec3", (0.0, 0.0, 0.0))}),
    SolverPBF: ({"particle_mask": ("int32", 0), "rest_density": ("float32", 1000.0),
                 "surface_tension": ("float32", 0.01), "viscosity": ("float32", 0.005)},
                {"rho": ("float32", 0.0), "pbf_lambda": ("float32", 0.0), "neighbor_count": ("int32", 0),
                 "delta_p": ("vec3", (0.0, 0.0, 0.0)), "step_displacement": ("vec3", (0.0, 0.0, 0.0))}),
}


@pytest.mark.parametrize("solver", [SolverWCSPH, SolverDFSPH, SolverPBF])
def test_custom_attributes_match_reference(solver):
    b = ModelBuilder(up_axis=Axis.Z)
    solver.register_custom_attributes(b)
    model_ref, state_ref = REF_ATTRS[solver]
    got_model = {a.name: (a.dtype, a.default) for a in b._attrs.values() if a.assignment == AttributeAssignment.MODEL}
    got_state = {a.name: (a.dtype, tuple(a.default) if isinstance(a.default, (list, tuple)) else a.default)
                 for a in b._attrs.values() if a.assignment == AttributeAssignment.STATE}
    assert got_model == model_ref
    assert got_state == state_ref
    assert all(a.namespace == "sph" for a in b._attrs.values())


def test_builder_model_state_on_cpu():
    b = ModelBuilder(up_axis=Axis.Z)
    b.default_particle_radius = 0.01
    SolverWCSPH.register_custom_attributes(b)
    pos = np.arange(12, dtype=np.float32).reshape(4, 3)
    b.add_particles(pos=pos, vel=np.zeros_like(pos), mass=[1.0] * 4, radius=[0.01] * 4,
                    custom_attributes={"sph:stiffness": [1.0, 2.0, 3.0, 4.0], "sph:particle_mask": [0, 1, 0, 0]})
    b.add_ground_plane()
    m = b.finalize(device="cpu")
    m.set_gravity((0.0, 0.0, -10.0))
    assert m.particle_count == 4
    np.testing.assert_array_equal(m.sph.stiffness.numpy(), [1, 2, 3, 4])
    np.testing.assert_array_equal(m.sph.exponent.numpy(), [7.0] * 4)         # default
    np.testing.assert_array_equal(m.sph.particle_mask.numpy(), [0, 1, 0, 0])
    np.testing.assert_array_equal(m.particle_flags.numpy(), [int(ParticleFlags.ACTIVE)] * 4)
    np.testing.assert_array_equal(m.gravity.numpy(), [[0, 0, -10]])
    assert m.particle_max_radius == pytest.approx(0.01)
    s0, s1 = m.state(), m.state()
    np.testing.assert_array_equal(s0.particle_q.numpy(), pos)
    s0.particle_q.assign(pos + 1)
    np.testing.assert_array_equal(s1.particle_q.numpy(), pos)                  # states are independent copies
    np.testing.assert_array_equal(s0.sph.rho.numpy(), np.zeros(4))
    assert m.particle_grid is not None and m.particle_grid.dims == (128, 128, 128)
    assert Model.AttributeFrequency.PARTICLE is not None


def test_unknown_attribute_and_bad_lengths_are_rejected():
    b = ModelBuilder()
    SolverPBF.register_custom_attributes(b)
    with pytest.raises(KeyError):
        b.add_particles(pos=[[0, 0, 0]], vel=[[0, 0, 0]], mass=[1.0], custom_attributes={"sph:nope": [1.0]})
    with pytest.raises(ValueError):
        b.add_particles(pos=[[0, 0, 0]], vel=[[0, 0, 0]], mass=[1.0], custom_attributes={"sph:viscosity": [1, 2]})


@pytest.mark.parametrize("method", ["wcsph", "dfsph", "pbf"])
def test_solvers_fail_loudly_without_a_gpu(method):
    from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

    with pytest.raises(RuntimeError, match="HIP device"):
        DamBreakSimulation(make_scene(512), method, device="cpu")


def test_driver_validates_wcsph_options():
    from fluid_simulation_amd.examples.dam_break.simulation import DamBreakSimulation

    for kw in ({"wcsph_stiffness": 0.0}, {"wcsph_viscosity": float("nan")}, {"wcsph_smoothing_length_coff": -1}):
        with pytest.raises(ValueError):
            DamBreakSimulation(make_scene(512), "wcsph", device="cpu", **kw)
    with pytest.raises(ValueError):
        DamBreakSimulation(make_scene(512), "apic", device="cpu")


# ---------------------------------------------------------------- C-ABI

def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(fs_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 18
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.SIGNATURES)
    assert lib.fs_version() >= 100
    assert lib.fs_grid_workspace_bytes(1000, 128, 128, 128) > 8 * 128 ** 3    # start + count per bucket
    assert lib.fs_work_bytes(1000, 96) > lib.fs_work_bytes(1000, 0) > 0


def test_capacity_above_the_addressing_limit_is_rejected():
    """Positions are read with 32-bit byte offsets: a grid or workspace for
    more than FS_MAX_PARTICLES (2^28 - 1) points is refused before any device
    work (the workspace pointer is never touched)."""
    lib = _lib.load()
    limit = (1 << 28) - 1
    assert f"#define FS_MAX_PARTICLES ((1 << 28) - 1)" in open(HEADER).read()
    g = _lib.FsGrid()
    assert lib.fs_grid_init(ctypes.byref(g), ctypes.c_void_p(256), 1 << 62, limit + 1, 128, 128, 128, 0, None) != 0
    assert b"maximum" in lib.fs_last_error()
    w = _lib.FsWork()
    assert lib.fs_work_init(ctypes.byref(w), ctypes.c_void_p(256), 1 << 62, limit + 1, 0, None) != 0
    assert b"maximum" in lib.fs_last_error()


def test_struct_layouts_agree_with_header(tmp_path):
    """Compile a C probe against include/fluidsim.h and compare sizeof/offsetof
    with the ctypes mirrors."""
    structs = {"fs_grid": _lib.FsGrid, "fs_model": _lib.FsModel, "fs_work": _lib.FsWork,
               "fs_wcsph_params": _lib.FsWcsphParams, "fs_wcsph_io": _lib.FsWcsphIO,
               "fs_dfsph_params": _lib.FsDfsphParams, "fs_dfsph_io": _lib.FsDfsphIO,
               "fs_pbf_params": _lib.FsPbfParams, "fs_pbf_io": _lib.FsPbfIO,
               "fs_scene": _lib.FsScene, "fs_batch_io": _lib.FsBatchIO,
               "fs_wcsph_adjoint_io": _lib.FsWcsphAdjointIO, "fs_separation_params": _lib.FsSeparationParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fluidsim.h"', "int main(void){"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            c, f, v = line.split()
            got[(c, f)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_bench_reads_the_committed_profiles_for_the_default_variants():
    """bench.py prices the dominant kernel's launch from the committed
    rocprofv3 summaries: HBM traffic, VALU instructions and cache hit rates
    of the default launch variants must be there."""
    sys.path.insert(0, ROOT)
    import bench
    from fluid_simulation_amd import autotune as at

    for pass_name, kernel in (("wcsph_density", "k_wcsph_density"), ("wcsph_force", "k_wcsph_force")):
        v0 = at.describe(pass_name)[0]
        key = bench.variant_key(kernel, {"block": v0["block"], "list_depth": v0["list_depth"], "unroll": v0["unroll"],
                                         "particles_per_thread": v0["particles_per_thread"],
                                         "lanes_per_particle": v0["lanes_per_particle"], "coop": v0["coop"]})
        assert bench.pmc_traffic(key) > 0, key
        assert bench.pmc_valu_instructions(key) > 0, key
        l1, l2 = bench.pmc_cache(key)
        assert 0.0 < l1 <= 1.0 and 0.0 < l2 <= 1.0, key
    assert bench.pmc_cache("k_no_such_kernel<1>") is None


def test_bench_refuses_multi_gpu():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 2
    assert "one GPU" in res.stderr


def test_tune_table_round_trip_and_cache(tmp_path, monkeypatch):
    """The launch-variant table is host state: describe/select/selected work
    without a GPU, unknown passes and variants are rejected, and the tuner's
    cache is keyed by device and workload."""
    import ctypes
    import json

    from fluid_simulation_amd import autotune as at

    lib = _lib.load()
    for name in at.PASSES:
        vs = at.variants(name)
        assert len(vs) >= 2 and all(b % 64 == 0 and d >= 16 for b, d in vs)
        assert at.selected(name) == 0
        at.select(name, len(vs) - 1)
        assert at.selected(name) == len(vs) - 1
        at.select(name, 0)
    assert lib.fs_tune_select(99, 0) != 0
    assert lib.fs_tune_select(0, 999) != 0
    assert lib.fs_tune_describe(0, -1, None, None) != 0
    b, d = ctypes.c_int32(), ctypes.c_int32()
    assert lib.fs_tune_describe(0, 0, ctypes.byref(b), ctypes.byref(d)) == 0 and (b.value, d.value) == at.variants(
        "wcsph_density")[0]
    # cached selection is applied without timing anything
    monkeypatch.setenv("FLUIDSIM_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(at, "device_arch", lambda: "dev")
    (tmp_path / "tune.json").write_text(json.dumps({"dev|w": {"wcsph_force": {"variant": 2, "block": 256,
                                                                             "list_depth": 16}}}))
    res = at.autotune(lambda: (_ for _ in ()).throw(AssertionError("must not run")), ("wcsph_force",), "w")
    assert res["wcsph_force"]["cached"] and at.selected("wcsph_force") == 2
    at.select("wcsph_force", 0)
    # every solver's passes are tunable; variants carry unroll and particles per lane
    assert set(p for ps in at.SOLVER_PASSES.values() for p in ps) <= set(at.PASSES)
    shapes = {(v["unroll"], v["particles_per_thread"]) for v in at.describe("wcsph_force")}
    assert {(4, 1), (8, 1), (4, 2)} <= shapes
    assert {v["unroll"] for v in at.describe("dfsph_replay")} == {2, 4, 8}
    assert {v["lanes_per_particle"] for v in at.describe("dfsph_replay")} == {1, 2, 4}
    assert all(v["block"] // v["lanes_per_particle"] >= 64 for p in at.PASSES for v in at.describe(p))
    assert at.workload_key("dfsph", 401856) == "dfsph-n524288" and at.workload_key("pbf", 512) == "pbf-n512"
