"""SIRT.SIRT().main -- futhark/SIRT.fut:9-49."""
from __future__ import annotations

from ._runtime import DeviceArray, EntryPoint, f32_array, f32_host, f32_scalar, i32_scalar, image_side


class SIRT(EntryPoint):
    def main(self, angles, rhozero, deltarho, image, projections, iterations) -> DeviceArray:
        """angles [A] f32, rhozero f32, deltarho f32, image [n] f32 (initial guess),
        projections [p] f32, iterations i32 -> reconstructed image [n] f32.

        x <- x + C*B(R*(b - A x)) with R = 1/rowsum, C = 1/colsum (0 where the sum is 0),
        no relaxation and no clamp.  The caller's image is not modified (the reference
        consumes only its device copy)."""
        a = f32_host(angles, 0)
        rz = f32_scalar(rhozero, 1)
        dr = f32_scalar(deltarho, 2)
        img = f32_array(image, 3, self.torch_device, copy=True)
        proj = f32_array(projections, 4, self.torch_device)
        iters = i32_scalar(iterations, 5)
        size = image_side(img.numel())
        if a.size == 0 or proj.numel() % a.size:
            raise ValueError(f"projections has {proj.numel()} elements, not a multiple of {a.size} angles")
        g = self.ctx.geometry(a, rz, dr, proj.numel() // a.size, size)
        if iters > 0:
            g.sirt_(img, proj, iters)
        return DeviceArray(img)
