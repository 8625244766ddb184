"""Autotuner for the launch variants of the tunable neighbour passes.

The library exposes, per tunable pass, a few launch variants: workgroup size
x depth of the per-lane LDS survivor list. The list depth sets how many
workgroups fit in a CU's LDS, and so the occupancy. All variants give
bit-identical results. ``tune`` times every variant of each requested pass on
the caller's own workload, using the library's per-kernel HIP-event timing,
and selects the fastest. ``autotune`` does the same behind a JSON cache keyed
by device and workload.
"""
from __future__ import annotations

import ctypes
import json
import os
import re

import torch

from . import _lib

PASSES = {"wcsph_density": 0, "wcsph_force": 1, "batch_density": 2, "batch_force": 3}
KERNEL = {"wcsph_density": "k_wcsph_density", "wcsph_force": "k_wcsph_force",
          "batch_density": "k_wcsph_density", "batch_force": "k_wcsph_force"}


def variants(pass_name: str) -> list[tuple[int, int]]:
    """(workgroup size, LDS list depth) of every variant of a pass."""
    lib = _lib.load()
    pid = PASSES[pass_name]
    out = []
    for v in range(lib.fs_tune_variant_count(pid)):
        b, d = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(lib.fs_tune_describe(pid, v, ctypes.byref(b), ctypes.byref(d)), "fs_tune_describe")
        out.append((b.value, d.value))
    return out


def select(pass_name: str, variant: int) -> None:
    _lib.check(_lib.load().fs_tune_select(PASSES[pass_name], int(variant)), "fs_tune_select")


def selected(pass_name: str) -> int:
    return int(_lib.load().fs_tune_selected(PASSES[pass_name]))


def _kernel_ms(name: str) -> float:
    total = 0.0
    for k, (_, ms) in _lib.profile_report().items():
        if re.sub(r"<.*>", "", k.strip("()")) == name:
            total += ms
    return total


def tune(step, passes, *, warmup: int = 2, reps: int = 5) -> dict:
    """Time ``step()`` (enqueues one substep on the current stream) under
    every variant of each pass in ``passes``, one pass at a time, and select
    the variant with the least time in that pass's kernel.  Returns
    {pass: {"variant", "block", "list_depth", "us_per_launch": [...]}}."""
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
        b, d = variants(name)[best]
        result[name] = {"variant": best, "block": b, "list_depth": d, "us_per_launch": times}
    lib.fs_profile_reset()
    return result


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
    """``tune`` behind a cache keyed by device name and ``workload``."""
    key = f"{torch.cuda.get_device_name()}|{workload}"
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
