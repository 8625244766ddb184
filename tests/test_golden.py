"""The FP32 oracle reproduces the committed golden fixtures bit for bit, and
the fixtures carry the reference's known answers (unit_test_dam_break.py:93-107
for the tetrahedron) and the structural facts of the Warp query."""
import numpy as np
import pytest

from tests.golden_cases import GRID_CASES, SUBSTEP_CASES, load, outputs


@pytest.mark.parametrize("case", GRID_CASES)
def test_grid_fixture_reproduced(case, oracle):
    fx = load(f"grid_{case}")
    pts, h = fx["points"], float(fx["radius"])
    g = oracle.OracleGrid().build(pts, h)
    np.testing.assert_array_equal(g.cell_of, fx["cell_of"])
    np.testing.assert_array_equal(g.point_ids, fx["point_ids"])
    nz = np.nonzero(g.cell_ends)[0]
    np.testing.assert_array_equal(nz, fx["buckets"])
    np.testing.assert_array_equal(g.cell_starts[nz], fx["bucket_starts"])
    off, idx = g.candidates(pts, h, None)
    np.testing.assert_array_equal(off, fx["cand_offsets"])
    np.testing.assert_array_equal(idx, fx["cand_indices"])
    # within-h sets are the candidates passing the support test, in visit order
    woff, widx = fx["within_offsets"], fx["within_indices"]
    hf = np.float32(h)
    for i in range(0, len(pts), max(1, len(pts) // 50)):
        cand = idx[off[i]:off[i + 1]]
        dd = (pts[i] - pts[cand]).astype(np.float32)        # float32 per operation, no contraction
        d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
        r = np.sqrt(d2.astype(np.float32))
        expect = cand[(r / hf) <= np.float32(1.0)]
        np.testing.assert_array_equal(widx[woff[i]:woff[i + 1]], expect)


@pytest.mark.parametrize("case", SUBSTEP_CASES)
def test_substep_fixture_reproduced(case, oracle):
    from tests.golden.make_golden import run_substep

    fx = load(f"substep_{case}")
    cfg_cls = {"wcsph": oracle.WcsphConfig, "dfsph": oracle.DfsphConfig, "pbf": oracle.PbfConfig}[fx["method"]]
    res = run_substep(fx["method"], fx["x"], fx["v"], fx["flags"], fx["mask"], float(fx["mass"]),
                      float(fx["radius"]), fx["material"], cfg_cls(**fx["config"]), float(fx["dt"]),
                      tuple(fx["gravity"]))
    for k, v in outputs(fx).items():
        np.testing.assert_array_equal(res[k], v, err_msg=f"{case}:{k}")


def test_tetrahedron_fixture_holds_the_reference_known_answer():
    """unit_test_dam_break.py:93-107: four compressed particles all over rest
    density and pressure, pushed apart, with zero net momentum."""
    fx = load("substep_wcsph_tetrahedron")
    assert np.all(fx["sph_rho"] > 1000.0)
    assert np.all(fx["sph_pressure"] > 0.0)
    off = fx["x"] - fx["x"].mean(axis=0)
    assert np.all(np.sum(fx["v_out"] * off, axis=1) > 0.0)
    np.testing.assert_allclose(fx["v_out"].sum(axis=0), 0.0, atol=1e-6)
