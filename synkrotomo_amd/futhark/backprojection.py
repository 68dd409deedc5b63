"""backprojection.backprojection().main -- futhark/backprojection.fut:82-93."""
from __future__ import annotations

from ._runtime import DeviceArray, EntryPoint, f32_array, f32_host, f32_scalar, i32_scalar


class backprojection(EntryPoint):
    def main(self, angles, rhozero, deltarho, size, projections) -> DeviceArray:
        """B.r: angles [A] f32, rhozero f32, deltarho f32, size i32, projections [p] f32
        (p = A*numrhos, angle-major) -> [size*size] f32."""
        a = f32_host(angles, 0)
        rz = f32_scalar(rhozero, 1)
        dr = f32_scalar(deltarho, 2)
        N = i32_scalar(size, 3)
        proj = f32_array(projections, 4, self.torch_device)
        if a.size == 0 or proj.numel() % a.size:
            raise ValueError(f"projections has {proj.numel()} elements, not a multiple of {a.size} angles")
        g = self.ctx.geometry(a, rz, dr, proj.numel() // a.size, N)
        return DeviceArray(g.bp(proj))
