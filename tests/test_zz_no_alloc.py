"""fs_batch_cfl_dt_noalloc (fixed device scratch, no allocation) against
fs_batch_cfl_dt (caller's scratch), through the C-ABI: the scene table, the
remaining frame time and the status must be bit-identical, on two calls in a
row, the second with fewer scenes and particles than the first."""
import ctypes

import numpy as np
import pytest
import torch

from fluid_simulation_amd import _lib

pytestmark = pytest.mark.gpu

FRAME_DT = 1.0 / 60.0


def _table(counts, rng):
    table = (_lib.FsScene * len(counts))()
    off = 0
    for k, c in enumerate(counts):
        t = table[k]
        t.particle_offset = off
        t.particle_count = int(c)
        h = rng.uniform(0.01, 0.03)
        t.smoothing_length = float(np.float32(h))
        t.cfl_smoothing_length = float(h)
        t.dt = -1.0
        off += int(c)
    return table, off


def _inputs(counts, seed, nonfinite=False):
    rng = np.random.default_rng(seed)
    table, n = _table(counts, rng)
    v = (rng.normal(size=(n, 3)) * rng.uniform(0.1, 3.0)).astype(np.float32)
    a = (rng.normal(size=(n, 3)) * 30.0).astype(np.float32)
    if nonfinite:
        v[n // 2, 1] = np.inf
    remaining = rng.uniform(0.0, FRAME_DT, len(counts))
    remaining[0] = FRAME_DT
    remaining[-1] = 0.0          # a finished scene: dt 0, remaining unchanged
    return table, v, a, remaining


def _run(noalloc, table, v, a, remaining, use_accel):
    dev = torch.device("cuda:0")
    lib = _lib.load()
    S = len(table)
    tab = torch.frombuffer(bytearray(bytes(table)), dtype=torch.uint8).to(dev)
    vt = torch.from_numpy(v).to(dev)
    at = torch.from_numpy(a).to(dev)
    rem = torch.from_numpy(remaining.copy()).to(dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    args = [ctypes.c_void_p(tab.data_ptr()), S, int(max(t.particle_count for t in table)),
            ctypes.c_void_p(vt.data_ptr()), ctypes.c_void_p(at.data_ptr()), use_accel, 250000.0, 7.0, 1000.0, 1.0e-6,
            FRAME_DT, 4, ctypes.c_void_p(rem.data_ptr())]
    if noalloc:
        rc = lib.fs_batch_cfl_dt_noalloc(*args, ctypes.c_void_p(status.data_ptr()), stream)
        _lib.check(rc, "fs_batch_cfl_dt_noalloc")
    else:
        scratch = torch.zeros(2 * S, dtype=torch.int32, device=dev)
        rc = lib.fs_batch_cfl_dt(*args, ctypes.c_void_p(scratch.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                 stream)
        _lib.check(rc, "fs_batch_cfl_dt")
    torch.cuda.synchronize()
    return tab.cpu().numpy(), rem.cpu().numpy(), int(status.item())


def test_noalloc_cfl_matches_the_scratch_form_on_two_calls():
    first = _inputs(np.random.default_rng(1).integers(1, 60000, 300), seed=2)
    second = _inputs(np.random.default_rng(3).integers(1, 5000, 37), seed=4)
    for k, (table, v, a, remaining) in enumerate((first, second)):
        for use_accel in (1, 0):
            ref = _run(False, table, v, a, remaining, use_accel)
            got = _run(True, table, v, a, remaining, use_accel)
            print(f"call {k}: {len(table)} scenes, {len(v)} particles, use_accel={use_accel}, status={got[2]}",
                  flush=True)
            np.testing.assert_array_equal(got[0], ref[0], err_msg=f"scene table, call {k}")
            np.testing.assert_array_equal(got[1].view(np.uint64), ref[1].view(np.uint64),
                                          err_msg=f"remaining, call {k}")
            assert got[2] == ref[2] == 0
    # the dt written for a live scene is positive, and a finished one stays at 0
    table, v, a, remaining = second
    got = _run(True, table, v, a, remaining, 1)
    out = (_lib.FsScene * len(table)).from_buffer_copy(got[0].tobytes())
    assert out[0].dt > 0.0 and out[len(table) - 1].dt == 0.0
    assert got[1][-1] == 0.0


def test_noalloc_cfl_flags_nonfinite_like_the_scratch_form():
    table, v, a, remaining = _inputs(np.random.default_rng(5).integers(100, 3000, 9), seed=6, nonfinite=True)
    ref = _run(False, table, v, a, remaining, 1)
    got = _run(True, table, v, a, remaining, 1)
    # non-finite maximum (1); its dt limit is then 0, below 1e-10 (2); the
    # status keeps the larger code
    assert ref[2] == got[2] == 2
    np.testing.assert_array_equal(got[0], ref[0])


def test_noalloc_cfl_rejects_more_scenes_than_its_scratch():
    lib = _lib.load()
    S = 4097
    dev = torch.device("cuda:0")
    table, _ = _table([1] * S, np.random.default_rng(0))
    tab = torch.frombuffer(bytearray(bytes(table)), dtype=torch.uint8).to(dev)
    v = torch.zeros((S, 3), dtype=torch.float32, device=dev)
    rem = torch.zeros(S, dtype=torch.float64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = lib.fs_batch_cfl_dt_noalloc(ctypes.c_void_p(tab.data_ptr()), S, 1, ctypes.c_void_p(v.data_ptr()),
                                     ctypes.c_void_p(v.data_ptr()), 1, 250000.0, 7.0, 1000.0, 0.0, FRAME_DT, 4,
                                     ctypes.c_void_p(rem.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                     ctypes.c_void_p(_lib.stream_handle(dev)))
    assert rc == _lib.FS_ERR_CAPACITY
    assert b"4096" in lib.fs_last_error()
