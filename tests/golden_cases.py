"""Loading the golden fixtures of tests/golden (written by
tests/golden/make_golden.py from the FP32 oracle)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRID_CASES = ["tetrahedron", "lattice5", "random_negative", "strip_aliasing", "dam_break_540"]
SUBSTEP_CASES = ["wcsph_tetrahedron", "wcsph_compressed", "dfsph_compressed", "pbf_compressed"]


def load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    for k in ("method", "material", "config"):
        if k in out:
            out[k] = str(out[k])
    if "material" in out:
        out["material"] = json.loads(out["material"])
        out["config"] = json.loads(out["config"])
    return out


def outputs(fx):
    return {k: v for k, v in fx.items() if k in ("x_out", "v_out", "accel") or k.startswith("sph_")}
