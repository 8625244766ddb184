"""The particle USD layer (host side).  pxr is not in this image, so the
layer is checked structurally against what the reference's SurfaceOutput
writes with --no-surface (examples/dam_break/surface.py:61-105, 130-165);
byte parity with a pxr-written layer is unpinned."""
import re

import numpy as np
import pytest

from fluid_simulation_amd.examples.dam_break.scene import make_scene
from fluid_simulation_amd.examples.dam_break.usd import (ParticleUsdWriter, height_guides, read_points_samples,
                                                         tank_edges)


def _write(tmp_path, frames, scene, name="out.usd"):
    w = ParticleUsdWriter(tmp_path / name, scene, "wcsph")
    rng = np.random.default_rng(0)
    xs = {}
    for f in frames:
        x = (scene.positions + rng.normal(size=scene.positions.shape) * 0.01).astype(np.float32)
        x[0] = [np.float32(1.0) / np.float32(3.0), -np.float32(1e-7), np.float32(123.456)]
        w.record(f, x)
        xs.setdefault(f, x)
    w.close()
    return w, xs


def test_layer_metadata_prims_and_samples(tmp_path):
    scene = make_scene(512)
    w, xs = _write(tmp_path, [0, 2, 4, 4, 6], scene)        # a repeated frame is recorded once
    text = (tmp_path / "out.usd").read_text()
    assert text.startswith("#usda 1.0\n")
    head = text[: text.index('def Xform "Fluid"')]
    for pat in (r'defaultPrim = "Fluid"', r'upAxis = "Z"', r"metersPerUnit = 1\b", r"startTimeCode = 0\b",
                r"endTimeCode = +6\n", r"framesPerSecond = 60\b", r"timeCodesPerSecond = 60\b",
                r'string solver = "wcsph"', r'string initial_position_sha256 = "%s"' % scene.report()[
                    "initial_position_sha256"]):
        assert re.search(pat, head), pat
    for prim in ('def Points "Particles"', 'def BasisCurves "TankGuide"', 'def "HeightGuides"',
                 'def BasisCurves "HeightScale"', 'def BasisCurves "RestVolume100"', 'def BasisCurves "RestVolume90"',
                 'def BasisCurves "RestVolume80"'):
        assert prim in text, prim
    assert text.count("{") == text.count("}") and text.count("(") == text.count(")")
    samples = read_points_samples(tmp_path / "out.usd")
    assert sorted(samples) == [0, 2, 4, 6]
    for f, x in xs.items():
        np.testing.assert_array_equal(samples[f], x)         # float32 values read back bit for bit
    assert w.samples == 4
    rep = w.report()
    assert rep["usd_samples"] == 4 and [s["frame"] for s in rep["frame_samples"]] == [0, 2, 4, 6]
    assert rep["enabled"] is False and rep["max_triangles"] == 0
    width = float(re.search(r'def Points "Particles"\n    \{[^}]*?float\[\] widths = \[([^\]]+)\]', text).group(1))
    assert width == pytest.approx(2.0 * scene.radius, rel=1e-7)


def test_tank_guide_is_the_box_edges(tmp_path):
    scene = make_scene(512)
    starts, ends = tank_edges(scene.spec)
    lo, hi = np.asarray(scene.spec.bound_lo, np.float32), np.asarray(scene.spec.bound_hi, np.float32)
    assert starts.shape == ends.shape == (12, 3)
    seg = np.abs(ends - starts)
    # every edge runs along one axis over the full box extent, and each
    # corner of the box meets three edges
    assert np.all((seg > 0).sum(axis=1) == 1)
    np.testing.assert_allclose(np.sort(seg.max(axis=1)), np.sort(np.repeat(hi - lo, 4)))
    corners, counts = np.unique(np.concatenate([starts, ends]), axis=0, return_counts=True)
    assert len(corners) == 8 and np.all(counts == 3)
    assert np.all((corners == lo) | (corners == hi))
    _write(tmp_path, [0], scene)
    text = (tmp_path / "out.usd").read_text()
    block = text[text.index('def BasisCurves "TankGuide"'):]
    block = block[: block.index("\n    }\n")]
    assert "int[] curveVertexCounts = [" + ", ".join(["2"] * 12) + "]" in block
    assert 'uniform token type = "linear"' in block and 'uniform token purpose = "guide"' in block


def test_height_guides_levels(tmp_path):
    scene = make_scene(512)
    guides = {g.name: g for g in height_guides(scene)}
    lo, hi = scene.spec.bound_lo, scene.spec.bound_hi
    ticks = guides["HeightScale"]
    assert len(ticks.starts) == 2 * len(np.arange(lo[2], hi[2] + 1e-6, 0.1))
    for frac in (1.0, 0.9, 0.8):
        g = guides[f"RestVolume{int(frac * 100)}"]
        z = np.float32(lo[2] + scene.ideal_height * frac)
        level = np.abs(g.starts[:, 2] - z) < 1e-6
        assert level.sum() >= 4 and g.volume_fraction == frac
    _write(tmp_path, [0], scene)
    text = (tmp_path / "out.usd").read_text()
    m = re.search(r'def BasisCurves "RestVolume90" \(\n\s+customData = \{\n\s+double restVolumeFraction = ([^\n]+)\n'
                  r'\s+double waterHeightMeters = ([^\n]+)\n', text)
    assert m and float(m.group(1)) == 0.9
    assert float(m.group(2)) == pytest.approx(0.9 * scene.ideal_height)


def test_refuses_binary_crate_and_nonfinite(tmp_path):
    scene = make_scene(512)
    with pytest.raises(ValueError):
        ParticleUsdWriter(tmp_path / "x.usdc", scene, "wcsph")
    w = ParticleUsdWriter(tmp_path / "x.usda", scene, "pbf")
    bad = scene.positions.copy()
    bad[3, 1] = np.nan
    with pytest.raises(RuntimeError):
        w.record(0, bad)
    w.close()
    assert read_points_samples(tmp_path / "x.usda") == {}
