"""A Warp-array-shaped view of a torch tensor.

The reference's solvers, driver and tests use Warp arrays (``.numpy()``,
``.assign()``, ``.zero_()``, ``wp.copy``).  ``Array`` keeps that surface over a
torch tensor so the same calling code works; ``.tensor`` is the tensor and
``.ptr`` its device address for the C-ABI.
"""
from __future__ import annotations

import numpy as np
import torch

# dtype tokens used by custom attributes (Warp: wp.float32, wp.int32, wp.vec3, int)
float32 = "float32"
int32 = "int32"
vec3 = "vec3"

_TORCH = {float32: torch.float32, int32: torch.int32, vec3: torch.float32}


def dtype_token(dtype) -> str:
    if dtype in (float32, "float", float, torch.float32, np.float32):
        return float32
    if dtype in (int32, "int", int, torch.int32, np.int32):
        return int32
    if dtype in (vec3, "vec3"):
        return vec3
    raise TypeError(f"unsupported attribute dtype {dtype!r}")


class Array:
    """1-D array of scalars or vec3 (shape (N,) or (N,3)) backed by a tensor."""

    __slots__ = ("tensor", "dtype")

    def __init__(self, tensor: torch.Tensor, dtype: str | None = None):
        if dtype is None:
            dtype = vec3 if tensor.dim() == 2 and tensor.shape[-1] == 3 else (
                int32 if tensor.dtype == torch.int32 else float32)
        self.tensor = tensor.contiguous()
        self.dtype = dtype

    @classmethod
    def from_values(cls, values, dtype, device, count=None):
        dtype = dtype_token(dtype)
        arr = np.asarray(values, dtype=np.int32 if dtype == int32 else np.float32)
        if dtype == vec3:
            arr = arr.reshape(-1, 3)
        if count is not None and arr.shape[0] != count:
            raise ValueError(f"expected {count} values, got {arr.shape[0]}")
        return cls(torch.from_numpy(np.ascontiguousarray(arr)).to(device), dtype)

    @classmethod
    def full(cls, count, value, dtype, device):
        dtype = dtype_token(dtype)
        shape = (count, 3) if dtype == vec3 else (count,)
        t = torch.empty(shape, dtype=_TORCH[dtype], device=device)
        t[...] = torch.as_tensor(value, dtype=_TORCH[dtype])
        return cls(t, dtype)

    @classmethod
    def empty(cls, count, dtype, device):
        dtype = dtype_token(dtype)
        shape = (count, 3) if dtype == vec3 else (count,)
        return cls(torch.empty(shape, dtype=_TORCH[dtype], device=device), dtype)

    @property
    def ptr(self) -> int:
        return self.tensor.data_ptr()

    @property
    def shape(self):
        return (self.tensor.shape[0],)

    @property
    def size(self) -> int:
        return int(self.tensor.shape[0])

    @property
    def device(self):
        return self.tensor.device

    def __len__(self):
        return self.size

    def numpy(self) -> np.ndarray:
        return self.tensor.detach().cpu().numpy().copy()

    def assign(self, values) -> "Array":
        if isinstance(values, Array):
            values = values.tensor
        src = torch.as_tensor(np.asarray(values) if not isinstance(values, torch.Tensor) else values)
        self.tensor.copy_(src.reshape(self.tensor.shape).to(self.tensor.dtype))
        return self

    def zero_(self) -> "Array":
        self.tensor.zero_()
        return self

    def fill_(self, value) -> "Array":
        self.tensor[...] = torch.as_tensor(value, dtype=self.tensor.dtype)
        return self

    def clone(self) -> "Array":
        return Array(self.tensor.clone(), self.dtype)

    def __repr__(self):
        return f"Array({self.dtype}, n={self.size}, device={self.device})"


def copy(dst: Array, src: Array) -> None:
    """wp.copy(dst, src): device-side copy on the current stream."""
    dst.tensor.copy_(src.tensor)
