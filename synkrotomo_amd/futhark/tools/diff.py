"""diff.diff().main -- futhark/tools/diff.fut:1-7."""
from __future__ import annotations

import numpy as np

from .._runtime import EntryPoint
from ._common import run


class diff(EntryPoint):
    def main(self, sane, test) -> np.int32:
        """Number of indices i with |sane - test| > 0.05, after the reference's filter (> 0)
        on the index list, which also drops index 0."""
        return np.int32(run(self, sane, test)[1])
