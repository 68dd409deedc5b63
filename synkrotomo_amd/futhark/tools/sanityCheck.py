"""sanityCheck.sanityCheck().main -- futhark/tools/sanityCheck.fut:1-4."""
from __future__ import annotations

import numpy as np

from .._runtime import EntryPoint
from ._common import run


class sanityCheck(EntryPoint):
    def main(self, sane, test) -> np.bool_:
        """True iff every |sane - test| < 0.05 (a NaN difference fails)."""
        return np.bool_(run(self, sane, test)[0])
