from __future__ import annotations

from ..._native import lib  # noqa: F401  (fail early if the library is missing)
from ...compare import REFERENCE_TOL, compare
from .._runtime import EntryPoint, f32_array


def run(ep: EntryPoint, sane, test):
    a = f32_array(sane, 0, ep.torch_device)
    b = f32_array(test, 1, ep.torch_device)
    return compare(a, b, REFERENCE_TOL, ep.ctx)
