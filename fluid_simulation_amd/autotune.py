"""Autotuner for the launch variants of the tunable neighbour passes.

The library exposes, per tunable pass, a few launch variants: workgroup size,
depth of the per-lane LDS survivor list (it sets how many workgroups fit in a
CU's LDS, and so the occupancy), the number of neighbours a list replay
fetches together and the particles per lane.  All variants give bit-identical
results.  ``tune`` times every variant of each requested pass on the caller's
own workload, using the library's per-kernel HIP-event timing, and selects
the fastest.  ``autotune`` does the same behind a JSON cache keyed by the GPU
architecture and the workload (``workload_key``: solver and particle-count
bucket).
"""
from __future__ import annotations

import ctypes
import json
import os
import re

import torch

from . import _lib

PASSES = {"wcsph_density": 0, "wcsph_force": 1, "batch_density": 2, "batch_force": 3, "dfsph_scan": 4,
          "dfsph_replay": 5, "pbf_scan": 6, "pbf_replay": 7}
# the kernels whose time a pass's variant changes
KERNEL = {"wcsph_density": ("k_wcsph_density", "k_wcsph_density_coop"), "wcsph_force": ("k_wcsph_force",),
          "batch_density": ("k_wcsph_density", "k_wcsph_density_coop"), "batch_force": ("k_wcsph_force",),
          "dfsph_scan": ("k_dfsph_density_factor",),
          "dfsph_replay": ("k_dfsph_nonpressure", "k_dfsph_rate", "k_dfsph_correction"),
          "pbf_scan": ("k_pbf_density", "k_pbf_lambda"),
          "pbf_replay": ("k_pbf_nonpressure", "k_pbf_delta", "k_pbf_xsph")}
# the passes each solver's substep can tune
SOLVER_PASSES = {"wcsph": ("wcsph_density", "wcsph_force"), "dfsph": ("dfsph_scan", "dfsph_replay"),
                 "pbf": ("pbf_scan", "pbf_replay")}


def describe(pass_name: str) -> list[dict]:
    """Every variant of a pass: workgroup size, LDS list depth (chunk slots of
    a cooperative scan), replay unroll, particles per lane, lanes per
    particle, and the cooperative scan's chunks per group (0: per-lane
    scan)."""
    lib = _lib.load()
    pid = PASSES[pass_name]
    out = []
    for v in range(lib.fs_tune_variant_count(pid)):
        b, d, u, p, l = (ctypes.c_int32() for _ in range(5))
        _lib.check(lib.fs_tune_describe_ex(pid, v, ctypes.byref(b), ctypes.byref(d), ctypes.byref(u),
                                           ctypes.byref(p), ctypes.byref(l)), "fs_tune_describe_ex")
        coop = lib.fs_tune_variant_coop(pid, v)
        if coop < 0:
            _lib.check(coop, "fs_tune_variant_coop")
        out.append({"block": b.value, "list_depth": d.value, "unroll": u.value, "particles_per_thread": p.value,
                    "lanes_per_particle": l.value, "coop": int(coop)})
    return out


def variants(pass_name: str) -> list[tuple[int, int]]:
    """(workgroup size, LDS list depth) of every variant of a pass."""
    return [(v["block"], v["list_depth"]) for v in describe(pass_name)]


def workload_key(solver: str, n: int) -> str:
    """Cache key of a workload: the solver and its particle count rounded up
    to a power of two."""
    return f"{solver}-n{1 << max(0, int(n) - 1).bit_length()}"


def select(pass_name: str, variant: int) -> None:
    _lib.check(_lib.load().fs_tune_select(PASSES[pass_name], int(variant)), "fs_tune_select")


def selected(pass_name: str) -> int:
    return int(_lib.load().fs_tune_selected(PASSES[pass_name]))


def _kernel_ms(names) -> float:
    total = 0.0
    for k, (_, ms) in _lib.profile_report().items():
        if re.sub(r"<.*>", "", k.strip("()")) in names:
            total += ms
    return total


def tune(step, passes, *, warmup: int = 2, reps: int = 5) -> dict:
    """Time ``step()`` (enqueues one substep on the current stream) under
    every variant of each pass in ``passes``, one pass at a time, and select
    the variant with the least time in that pass's kernels.  Returns
    {pass: {"variant", "block", "list_depth", "unroll", "particles_per_thread",
    "us_per_launch": [...]}} (us_per_launch: the pass's kernel time per
    step, per variant)."""
    lib = _lib.load()
    result = {}
    for name in passes:
        times = []
        for v in range(len(variants(name))):
            select(name, v)
            for _ in range(warmup):
                step()
            torch.cuda.synchronize()
            lib.fs_profile_reset()
            lib.fs_profile_enable(1)
            for _ in range(reps):
                step()
            lib.fs_profile_enable(0)
            times.append(1e3 * _kernel_ms(KERNEL[name]) / reps)
        best = min(range(len(times)), key=times.__getitem__)
        select(name, best)
        result[name] = dict(variant=best, **describe(name)[best], us_per_launch=times)
    lib.fs_profile_reset()
    return result


def device_arch() -> str:
    """The GPU architecture name (e.g. gfx950...), the cache's device key."""
    return torch.cuda.get_device_properties(0).gcnArchName


def cache_path() -> str:
    return os.environ.get("FLUIDSIM_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "fluid_simulation_amd", "tune.json"))


def _load_cache() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(step, passes, workload: str, *, use_cache: bool = True, **kw) -> dict:
    """``tune`` behind a cache keyed by the GPU architecture and ``workload``
    (e.g. ``workload_key(solver, n)``)."""
    key = f"{device_arch()}|{workload}"
    cache = _load_cache() if use_cache else {}
    hit = cache.get(key)
    if hit and all(p in hit for p in passes):
        for p in passes:
            select(p, hit[p]["variant"])
        return {p: dict(hit[p], cached=True) for p in passes}
    result = tune(step, passes, **kw)
    if use_cache:
        cache[key] = result
        path = cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return result
