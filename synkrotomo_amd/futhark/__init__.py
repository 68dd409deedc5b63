"""Reference-compatible entry points (tomograph/synkrotomo's ``futhark`` package).

    from synkrotomo_amd.futhark import SIRT
    result = SIRT.SIRT().main(theta.astype(np.float32), rhozero, deltarho, image,
                              sinogram.flatten().astype(np.float32), iterations).get()

Modules and classes keep the reference's names and ``main`` argument orders
(README.md:21-25, Makefile:1-4); each call runs hand-written HIP through libsynkro.so.
"""
from . import (SIRT, SIRT3D, backprojection, backprojection_sparse,  # noqa: F401
               forwardprojection, forwardprojection_sparse, tools)

__all__ = ["SIRT", "SIRT3D", "backprojection", "backprojection_sparse", "forwardprojection",
           "forwardprojection_sparse", "tools"]
