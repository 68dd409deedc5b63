"""mse.mse().main -- futhark/tools/mse.fut:1-3."""
from __future__ import annotations

import numpy as np

from .._runtime import EntryPoint
from ._common import run


class mse(EntryPoint):
    def main(self, sane, test) -> np.float32:
        """sum((sane - test)**2) / n, the squares in f32 and their sum in FP64."""
        return np.float32(run(self, sane, test)[2])
