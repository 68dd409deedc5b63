"""Kernel-variant autotuning with a persistent cache.

``tuned(geom, batch)`` applies the cached choice for the geometry's shape (image size, angle
count, bins, bin width, device architecture) and the calls' batch size (rounded up to a power
of two, at most 64) or runs ``sk_autotune_batch`` on the GPU and stores the result: batches of
small images fill the chip where one image does not, and favour other kernel shapes.  The cache is a JSON file at $SYNKRO_TUNE_CACHE, default
~/.cache/synkrotomo_amd/tune.json.  Variants only change tile shapes and rays per lane --
every variant computes identical results -- so a stale cache can only cost time.
"""
from __future__ import annotations

import json
import os

import torch

from .projector import Geometry


def cache_path() -> str:
    return os.environ.get("SYNKRO_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "synkrotomo_amd", "tune.json"))


# Bumped whenever the meaning of the variant numbers changes, so old cache entries are ignored.
VARIANT_TABLE = 5


def batch_bucket(batch: int) -> int:
    """The batch size variants are tuned at for calls of `batch` images."""
    b = 1
    while b < min(int(batch), 64):
        b *= 2
    return b


def _key(geom: Geometry, batch: int = 1) -> str:
    arch = torch.cuda.get_device_properties(geom.ctx.device).gcnArchName.split(":")[0]
    return f"v{VARIANT_TABLE}/{arch}/N{geom.N}/A{geom.A}/R{geom.R}/d{geom.drho:g}/b{batch_bucket(batch)}"


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # read-only home: tuning still applies to this process


def tuned(geom: Geometry, retune: bool = False, batch: int = 1) -> tuple[int, int]:
    """Apply (and if needed measure) the fastest kernel variants for this geometry on calls of
    `batch` images."""
    table = _load()
    key = _key(geom, batch)
    if not retune and key in table:
        fv, bv = table[key]
        try:
            geom.set_variants(fv, bv)
            return fv, bv
        except ValueError:
            pass  # an entry the library no longer accepts: measure again
    fv, bv = geom.autotune(batch_bucket(batch))
    table[key] = [fv, bv]
    _store(table)
    return fv, bv
