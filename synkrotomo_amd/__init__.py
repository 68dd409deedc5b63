"""synkrotomo_amd -- MI355X-native projector and SIRT (drop-in for tomograph/synkrotomo).

Layers:
  libsynkro.so (csrc/*.hip)  hand-written HIP kernels behind a C ABI (include/synkro.h)
  projector                  contexts, geometries, operators on torch tensors
  futhark                    the reference's module/class/``main`` surface
"""
from .projector import Context, Geometry, Matrix, centred_bins, sirt_batched, uniform_angles  # noqa: F401

__version__ = "0.1.0"
