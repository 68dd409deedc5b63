"""SIRT3D.SIRT3D().main -- futhark/SIRT3D.fut:18-28.

The reference passes its ``rhos`` array where SIRT expects (rhozero, deltarho) and does not
type-check (SIRT3D.fut:27 vs SIRT.fut:9-14).  The evident intent is implemented: for
every slice i < size, SIRT from a zero image on projections[i*r*a:(i+1)*r*a] with
rhozero = rhos[0], deltarho = rhos[1] - rhos[0] and numrhos = len(rhos) (as
archive/projection_lib.fut:114-115 derives them); the output is the slices
concatenated, [size*size*size] f32.  All slices run as one batched launch sequence.
"""
from __future__ import annotations

import numpy as np
import torch

from ._runtime import DeviceArray, EntryPoint, f32_array, f32_host, i32_scalar


class SIRT3D(EntryPoint):
    def main(self, angles, rhos, projections, iterations, size) -> DeviceArray:
        a = f32_host(angles, 0)
        rh = f32_host(rhos, 1)
        proj = f32_array(projections, 2, self.torch_device)
        iters = i32_scalar(iterations, 3)
        N = i32_scalar(size, 4)
        if rh.size < 2:
            raise ValueError("rhos needs at least two entries")
        if N < 2 or N % 2:
            raise ValueError(f"size must be even and >= 2 (got {N})")
        R = rh.size
        if proj.numel() != N * a.size * R:
            raise ValueError(f"projections has {proj.numel()} elements, expected size*a*r = {N * a.size * R}")
        rz = float(rh[0])
        dr = float(np.float32(rh[1] - rh[0]))
        g = self.ctx.geometry(a, rz, dr, R, N)
        img = torch.zeros((N, N, N), dtype=torch.float32, device=self.torch_device)
        if iters > 0:
            g.sirt_(img, proj, iters)
        return DeviceArray(img.reshape(-1))
