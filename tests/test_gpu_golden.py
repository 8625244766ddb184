"""The HIP path against the committed golden fixtures: grid arrays and
candidate lists bit-exact, one solver substep within 1e-5 relative (integers
exact)."""
import numpy as np
import pytest
import torch

from fluid_simulation_amd.examples.dam_break.simulation import SOLVERS
from fluid_simulation_amd.hashgrid import HashGrid
from tests.cases import build_model, make_solver, rel_err
from tests.golden_cases import GRID_CASES, SUBSTEP_CASES, load, outputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", GRID_CASES)
def test_grid_matches_fixture(case):
    fx = load(f"grid_{case}")
    pts, h = fx["points"], float(fx["radius"])
    print(f"golden grid {case} n={len(pts)}", flush=True)
    g = HashGrid(128, 128, 128, device="cuda:0")
    t = torch.from_numpy(pts).cuda()
    g.reserve(len(pts))
    g.build(t, h)
    np.testing.assert_array_equal(g.cell_of.cpu().numpy(), fx["cell_of"])
    np.testing.assert_array_equal(g.point_ids.cpu().numpy(), fx["point_ids"])
    starts, ends = g.cell_tables()
    nz = fx["buckets"]
    np.testing.assert_array_equal(np.nonzero(ends.cpu().numpy())[0], nz)
    np.testing.assert_array_equal(starts.cpu().numpy()[nz], fx["bucket_starts"])
    np.testing.assert_array_equal(ends.cpu().numpy()[nz], fx["bucket_ends"])
    off, idx = g.query_lists(t, h, None)
    np.testing.assert_array_equal(off.cpu().numpy(), fx["cand_offsets"])
    np.testing.assert_array_equal(idx.cpu().numpy(), fx["cand_indices"])
    off, idx = g.query_lists(t, h, "cubic")
    np.testing.assert_array_equal(off.cpu().numpy(), fx["within_offsets"])
    np.testing.assert_array_equal(idx.cpu().numpy(), fx["within_indices"])


@pytest.mark.parametrize("case", SUBSTEP_CASES)
def test_substep_matches_fixture(case):
    fx = load(f"substep_{case}")
    method = fx["method"]
    print(f"golden substep {case}", flush=True)
    model = build_model(method, fx["x"], fx["v"], mass=float(fx["mass"]), radius=float(fx["radius"]),
                        mat=fx["material"], flags=fx["flags"], mask=fx["mask"], gravity=tuple(fx["gravity"]))
    cls = SOLVERS[method]
    cfg = (cls.PbfConfig if method == "pbf" else cls.SphConfig)(**fx["config"])
    solver = make_solver(method, model, cfg)
    s_in, s_out = model.state(), model.state()
    solver.step(s_in, s_out, None, None, float(fx["dt"]))
    torch.cuda.synchronize()
    got = {"x_out": s_out.particle_q.numpy(), "v_out": s_out.particle_qd.numpy(),
           "accel": solver._particle_accel.numpy()}
    for k in s_out.sph.names():
        got[f"sph_{k}"] = getattr(s_out.sph, k).numpy()
    for k, ref in outputs(fx).items():
        g = got[k]
        if ref.dtype.kind == "i":
            np.testing.assert_array_equal(g, ref, err_msg=f"{case}:{k}")
            continue
        floor = max(1e-12, 1e-3 * float(np.abs(ref).max(initial=0.0)))
        err = rel_err(g, ref, floor)
        exact = float(np.mean(g.astype(np.float32).view(np.uint32) == ref.astype(np.float32).view(np.uint32)))
        print(f"  {k:18s} exact={exact:.4f} rel_err={err:.2e}", flush=True)
        assert err <= 1e-5, f"{case}:{k}"
