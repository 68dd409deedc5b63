"""Argument handling and result type shared by the reference-compatible entry points.

The reference's generated PyOpenCL wrappers take 1-D float32 arrays and f32/i32 scalars,
raise TypeError for values of the wrong kind, copy inputs to the device and return a
device array whose ``.get()`` yields numpy (README.md:21-25).  ``DeviceArray`` plays the
role of that device array here, backed by a torch tensor in HBM.
"""
from __future__ import annotations

import numbers

import numpy as np
import torch

from ..projector import Context


class DeviceArray:
    """A 1-D float32 array resident on the GPU; ``.get()`` copies it to a numpy array."""

    __slots__ = ("tensor",)

    def __init__(self, tensor: torch.Tensor):
        self.tensor = tensor

    def get(self) -> np.ndarray:
        return self.tensor.detach().cpu().numpy()

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return np.dtype(np.float32)

    @property
    def size(self) -> int:
        return self.tensor.numel()

    def __len__(self) -> int:
        return self.tensor.shape[0]

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, device={self.tensor.device})"


def _bad(argno: int, futhark_type: str, value) -> TypeError:
    return TypeError(f"Argument #{argno} has invalid value\nFuthark type: {futhark_type}\n"
                     f"Argument has Python type {type(value)} and value: {value!r:.80}")


def f32_array(value, argno: int, device: torch.device, copy: bool = False) -> torch.Tensor:
    """A 1-D float32 argument ([]f32) as a contiguous device tensor."""
    if isinstance(value, DeviceArray):
        value = value.tensor
    if isinstance(value, torch.Tensor):
        if value.dtype != torch.float32 or value.dim() != 1:
            raise _bad(argno, "[]f32", value)
        t = value.to(device)
        t = t.contiguous()
        return t.clone() if (copy and t.data_ptr() == value.data_ptr()) else t
    if isinstance(value, np.ndarray):
        if value.dtype != np.float32 or value.ndim != 1:
            raise _bad(argno, "[]f32", value)
        return torch.from_numpy(np.ascontiguousarray(value)).to(device)
    raise _bad(argno, "[]f32", value)


def f32_host(value, argno: int) -> np.ndarray:
    """A 1-D float32 argument needed on the host (angles, rhos)."""
    if isinstance(value, DeviceArray):
        value = value.tensor
    if isinstance(value, torch.Tensor):
        if value.dtype != torch.float32 or value.dim() != 1:
            raise _bad(argno, "[]f32", value)
        return value.detach().cpu().numpy()
    if isinstance(value, np.ndarray):
        if value.dtype != np.float32 or value.ndim != 1:
            raise _bad(argno, "[]f32", value)
        return np.ascontiguousarray(value)
    raise _bad(argno, "[]f32", value)


def f32_scalar(value, argno: int) -> float:
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (numbers.Real, np.floating, np.integer)):
        raise _bad(argno, "f32", value)
    return float(np.float32(value))


def i32_scalar(value, argno: int) -> int:
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (numbers.Integral, np.integer)):
        raise _bad(argno, "i32", value)
    v = int(value)
    if not -(2 ** 31) <= v < 2 ** 31:
        raise _bad(argno, "i32", value)
    return v


def image_side(n: int) -> int:
    """size = t32(sqrt(n)) as the reference derives it (forwardprojection.fut:76,
    SIRT.fut:16); n must be the square of an even size."""
    size = int(np.float32(np.sqrt(np.float32(n))))
    if size * size != n:
        raise ValueError(f"image has {n} elements, which is not a square")
    if size % 2:
        raise ValueError(f"image side {size} is odd; the projector needs an even size")
    return size


class EntryPoint:
    """Base of the generated-class stand-ins: holds the device context."""

    def __init__(self, device: int = 0, **_options):
        # The generated classes accept OpenCL platform/device/tuning options; those have no
        # meaning here and are ignored.
        self.device = int(device)
        self.torch_device = torch.device("cuda", self.device)
        self._ctx = None

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = Context.default(self.device)
        return self._ctx
