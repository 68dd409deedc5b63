"""backprojection_sparse.backprojection_sparse().main --
futhark/backprojection_sparse.fut:26-37.

Same signature and result as backprojection.main; computed as a CSR SpMV with the
backprojection matrix B (rows = pixels) built on the GPU and cached per geometry."""
from __future__ import annotations

from ._runtime import DeviceArray, EntryPoint, f32_array, f32_host, f32_scalar, i32_scalar


class backprojection_sparse(EntryPoint):
    def main(self, angles, rhozero, deltarho, size, projections) -> DeviceArray:
        a = f32_host(angles, 0)
        rz = f32_scalar(rhozero, 1)
        dr = f32_scalar(deltarho, 2)
        N = i32_scalar(size, 3)
        proj = f32_array(projections, 4, self.torch_device)
        if a.size == 0 or proj.numel() % a.size:
            raise ValueError(f"projections has {proj.numel()} elements, not a multiple of {a.size} angles")
        if N < 2 or N % 2:
            raise ValueError(f"size must be even and >= 2 (got {N})")
        g = self.ctx.geometry(a, rz, dr, proj.numel() // a.size, N)
        return DeviceArray(g.sparse("bp").spmv(proj))
