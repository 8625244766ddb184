"""Time-sampled particle output as a USD text layer (USDA), without ``pxr``.

Counterpart of the particle path of the reference's ``SurfaceOutput``
(``examples/dam_break/surface.py:61-105, 130-165``) and of its static guides
(``examples/dam_break/guides.py``).  The layer holds:

* layer metadata: up axis Z, metres per unit 1, time codes and frames per
  second = the scene's fps, start time 0, end time = the last recorded frame,
  default prim ``/Fluid``, and ``customLayerData`` with the solver, the scene
  name and the SHA-256 of the initial positions (``surface.py:64-76``);
* ``/Fluid/Particles``: ``Points`` with constant width 2 r, display colour
  (0.25, 0.55, 0.95) and one time sample of ``points`` per recorded frame
  (``surface.py:82-86, 156-160``);
* ``/Fluid/TankGuide``: linear ``BasisCurves`` of the tank's 12 edges
  (``surface.py:12-22, 87-94``);
* ``/Fluid/HeightGuides/<name>``: the height ticks and the 100 / 90 / 80 %
  rest-volume levels with their labels (``guides.py:52-85``,
  ``surface.py:95-105``).

The surface mesh of the reference (Newton's ``ParticleSurface``) is not
produced.  USD's ``.usd`` extension accepts text layers, so ``.usd`` and
``.usda`` paths are written; ``.usdc`` (the binary crate format) is not.

The points block is streamed: samples are appended as frames are recorded
and the layer is completed by ``close()``, so a long run never holds more
than one frame of text in memory.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

PARTICLE_COLOR = (0.25, 0.55, 0.95)
TANK_COLOR = (0.55, 0.55, 0.60)


@dataclass
class Guide:
    name: str
    starts: np.ndarray
    ends: np.ndarray
    color: tuple
    width: float
    volume_fraction: float | None = None


def tank_edges(spec):
    """Start and end points of the 12 edges of the tank box (float32)."""
    lo, hi = np.asarray(spec.bound_lo, np.float64), np.asarray(spec.bound_hi, np.float64)
    bits = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    corners = np.where(bits == 1, hi, lo).astype(np.float32)
    bottom = [(k, (k + 1) % 4) for k in range(4)]
    edges = bottom + [(a + 4, b + 4) for a, b in bottom] + [(k, k + 4) for k in range(4)]
    a, b = zip(*edges)
    return corners[list(a)], corners[list(b)]


# seven-segment strokes of the label digits (guides.py:18-23): stroke k joins
# two of the six nodes of a 0.55 x 1 glyph box
_NODES = ((0.0, 1.0), (0.55, 1.0), (0.55, 0.5), (0.55, 0.0), (0.0, 0.0), (0.0, 0.5))
_STROKES = ((0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (5, 2))
_DIGITS = {"0": (0, 1, 2, 3, 4, 5), "1": (1, 2), "2": (0, 1, 6, 4, 3), "3": (0, 1, 2, 3, 6),
           "4": (5, 6, 1, 2), "5": (0, 5, 6, 2, 3), "6": (0, 5, 4, 3, 2, 6), "7": (0, 1, 2),
           "8": (0, 1, 2, 3, 4, 5, 6), "9": (0, 1, 2, 3, 5, 6)}


def _label(text, x, y, z, size):
    """Line segments spelling ``text`` (digits and '%') in the xz plane."""
    out = []
    for ch in text:
        if ch in _DIGITS:
            for k in _DIGITS[ch]:
                a, b = (_NODES[n] for n in _STROKES[k])
                out.append(((x + a[0] * size, y, z + a[1] * size), (x + b[0] * size, y, z + b[1] * size)))
        elif ch == "%":
            out.append(((x, y, z), (x + 0.6 * size, y, z + size)))
            for ox, oz in ((0.05, 0.8), (0.45, 0.05)):
                sq = ((ox, oz), (ox + 0.15, oz), (ox + 0.15, oz + 0.15), (ox, oz + 0.15))
                for k in range(4):
                    a, b = sq[k], sq[(k + 1) % 4]
                    out.append(((x + a[0] * size, y, z + a[1] * size), (x + b[0] * size, y, z + b[1] * size)))
        x += 0.85 * size
    return out


def height_guides(scene) -> list[Guide]:
    """Height ticks every 0.1 m on the two front posts, and the rest-volume
    water levels at 100 % (a closed loop) and 90 % / 80 % (dashed), each with
    a percentage label (guides.py:52-85)."""
    lo, hi = scene.spec.bound_lo, scene.spec.bound_hi
    y = lo[1] - 0.012
    ticks = [((x - 0.025, y, z), (x + 0.025, y, z))
             for z in np.arange(lo[2], hi[2] + 1.0e-6, 0.1) for x in (lo[0], hi[0])]
    guides = [Guide("HeightScale", np.asarray([a for a, _ in ticks], np.float32),
                    np.asarray([b for _, b in ticks], np.float32), (0.65, 0.65, 0.70), 0.002)]
    for frac, color in ((1.0, (1.0, 0.35, 0.10)), (0.9, (1.0, 0.70, 0.15)), (0.8, (1.0, 0.90, 0.35))):
        z = lo[2] + scene.ideal_height * frac
        if frac == 1.0:
            loop = [(lo[0], y, z), (hi[0], y, z), (hi[0], hi[1], z), (lo[0], hi[1], z)]
            segs = [(loop[k], loop[(k + 1) % 4]) for k in range(4)]
        else:
            segs = [((x, y, z), (min(x + 0.06, hi[0]), y, z)) for x in np.arange(lo[0], hi[0], 0.12)]
        size = min(0.026, scene.ideal_height * 0.075)
        segs += _label(f"{int(frac * 100)}%", hi[0] + 0.045, y, z - 0.5 * size, size)
        guides.append(Guide(f"RestVolume{int(frac * 100)}", np.asarray([a for a, _ in segs], np.float32),
                            np.asarray([b for _, b in segs], np.float32), color,
                            0.003 if frac == 1.0 else 0.0015, frac))
    return guides


def _f(v) -> str:
    """A float32 value, printed so that it reads back bit for bit."""
    return format(float(np.float32(v)), ".9g")


def _vec3_list(a) -> str:
    a = np.asarray(a, np.float32).reshape(-1, 3)
    return "[" + ", ".join(f"({_f(x)}, {_f(y)}, {_f(z)})" for x, y, z in a.tolist()) + "]"


def _color(c) -> str:
    return f"[({_f(c[0])}, {_f(c[1])}, {_f(c[2])})]"


def _string(s) -> str:
    return '"' + str(s).replace("\\", "\\\\").replace('"', '\\"') + '"'


class ParticleUsdWriter:
    """Streams particle positions into a USDA layer, one time sample per
    recorded frame (``record``), and completes it on ``close``."""

    _END_FIELD = 20   # fixed width of the endTimeCode value, rewritten on close

    def __init__(self, path, scene, method: str, scene_name: str = "APIC dam-break specifications"):
        path = Path(path)
        if path.suffix.lower() not in (".usd", ".usda"):
            raise ValueError("USD output is written as a text layer: use a .usd or .usda path (.usdc is the "
                             "binary crate format)")
        self.path = path.resolve()
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.scene = scene
        self.samples = 0
        self.frames: list[int] = []
        self._closed = False
        fps = float(scene.spec.fps)
        sha = scene.report()["initial_position_sha256"]
        self._fh = open(self.path, "wb")
        w = self._write
        w("#usda 1.0\n(\n")
        w("    customLayerData = {\n")
        w(f"        string initial_position_sha256 = {_string(sha)}\n")
        w(f"        string scene = {_string(scene_name)}\n")
        w(f"        string solver = {_string(method)}\n")
        w("    }\n")
        w('    defaultPrim = "Fluid"\n')
        self._end_offset = self._fh.tell() + len("    endTimeCode = ")
        w("    endTimeCode = " + "0".rjust(self._END_FIELD) + "\n")
        w(f"    framesPerSecond = {_f(fps)}\n")
        w("    metersPerUnit = 1\n")
        w("    startTimeCode = 0\n")
        w(f"    timeCodesPerSecond = {_f(fps)}\n")
        w('    upAxis = "Z"\n')
        w(")\n\n")
        w('def Xform "Fluid"\n{\n')
        starts, ends = tank_edges(scene.spec)
        self._curves("TankGuide", starts, ends, 0.004, TANK_COLOR, indent=1)
        w('    def "HeightGuides"\n    {\n')
        for g in height_guides(scene):
            extra = None
            if g.volume_fraction is not None:
                extra = {"restVolumeFraction": float(g.volume_fraction),
                         "waterHeightMeters": float(scene.ideal_height * g.volume_fraction)}
            self._curves(g.name, g.starts, g.ends, g.width, g.color, indent=2, custom=extra)
        w("    }\n\n")
        w('    def Points "Particles"\n    {\n')
        w(f"        color3f[] primvars:displayColor = {_color(PARTICLE_COLOR)}\n")
        w(f"        float[] widths = [{_f(scene.radius * 2.0)}] (\n            interpolation = \"constant\"\n"
          "        )\n")
        w("        point3f[] points.timeSamples = {\n")

    def _write(self, text: str) -> None:
        self._fh.write(text.encode("ascii"))

    def _curves(self, name, starts, ends, width, color, indent, custom=None):
        p = "    " * indent
        w = self._write
        pts = np.stack([np.asarray(starts, np.float32), np.asarray(ends, np.float32)], axis=1).reshape(-1, 3)
        w(f'{p}def BasisCurves "{name}"')
        if custom:
            w(f" (\n{p}    customData = {{\n")
            for k, v in custom.items():
                w(f"{p}        double {k} = {float(v)!r}\n")
            w(f"{p}    }}\n{p})")
        w(f"\n{p}{{\n")
        w(f"{p}    int[] curveVertexCounts = [{', '.join(['2'] * (len(pts) // 2))}]\n")
        w(f"{p}    point3f[] points = {_vec3_list(pts)}\n")
        w(f"{p}    color3f[] primvars:displayColor = {_color(color)}\n")
        w(f'{p}    uniform token purpose = "guide"\n')
        w(f'{p}    uniform token type = "linear"\n')
        w(f"{p}    float[] widths = [{_f(width)}] (\n{p}        interpolation = \"constant\"\n{p}    )\n")
        w(f"{p}}}\n\n")

    def record(self, frame: int, positions) -> None:
        """One time sample of the particle positions at time code ``frame``
        (a frame already recorded is skipped, as the reference skips it)."""
        if self._closed:
            raise RuntimeError("the USD layer is closed")
        frame = int(frame)
        if self.frames and self.frames[-1] == frame:
            return
        x = np.asarray(positions, np.float32)
        if not np.isfinite(x).all():
            raise RuntimeError("Cannot export non-finite particle positions")
        self._write(f"            {frame}: {_vec3_list(x)},\n")
        self.frames.append(frame)
        self.samples += 1

    def close(self) -> None:
        if self._closed:
            return
        self._write("        }\n    }\n}\n")
        end = str(self.frames[-1] if self.frames else 0)
        self._fh.flush()
        self._fh.seek(self._end_offset)
        self._write(end.rjust(self._END_FIELD))
        self._fh.close()
        self._closed = True

    def report(self) -> dict:
        """The surface record of run.json (surface.py:167-171), with no mesh."""
        return dict(enabled=False, parameters={}, usd_path=str(self.path), usd_samples=self.samples,
                    max_triangles=0, frame_samples=[dict(frame=f, triangles=0) for f in self.frames])


def read_points_samples(path) -> dict:
    """Time samples of /Fluid/Particles.points of a layer this module wrote:
    {frame: float32 (N, 3)} (a reader of this writer's own layout, for checks)."""
    text = Path(path).read_text()
    block = text[text.index("point3f[] points.timeSamples = {") :]
    block = block[block.index("{") + 1 : block.index("\n        }\n")]
    out = {}
    for line in block.strip().splitlines():
        key, _, rest = line.strip().partition(":")
        nums = rest.replace("(", " ").replace(")", " ").replace("[", " ").replace("]", " ").replace(",", " ").split()
        out[int(key)] = np.asarray([float(v) for v in nums], np.float64).astype(np.float32).reshape(-1, 3)
    return out


__all__ = ["ParticleUsdWriter", "read_points_samples", "tank_edges", "height_guides", "Guide"]
