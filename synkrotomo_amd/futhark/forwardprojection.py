"""forwardprojection.forwardprojection().main -- futhark/forwardprojection.fut:71-84."""
from __future__ import annotations

from ._runtime import DeviceArray, EntryPoint, f32_array, f32_host, f32_scalar, i32_scalar, image_side


class forwardprojection(EntryPoint):
    def main(self, angles, rhozero, deltarho, numrhos, image) -> DeviceArray:
        """A.x: angles [A] f32, rhozero f32, deltarho f32, numrhos i32, image [n] f32
        (n = size*size, row-major) -> [A*numrhos] f32, angle-major."""
        a = f32_host(angles, 0)
        rz = f32_scalar(rhozero, 1)
        dr = f32_scalar(deltarho, 2)
        R = i32_scalar(numrhos, 3)
        img = f32_array(image, 4, self.torch_device)
        size = image_side(img.numel())
        g = self.ctx.geometry(a, rz, dr, R, size)
        return DeviceArray(g.fp(img))
