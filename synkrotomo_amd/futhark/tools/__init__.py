"""The reference's comparison tools (futhark/tools/*.fut): sanityCheck, diff, mse."""
from . import diff, mse, sanityCheck  # noqa: F401

__all__ = ["diff", "mse", "sanityCheck"]
