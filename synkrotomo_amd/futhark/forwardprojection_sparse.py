"""forwardprojection_sparse.forwardprojection_sparse().main --
futhark/forwardprojection_sparse.fut:26-40.

Same signature and result as forwardprojection.main; computed as a CSR SpMV with the system
matrix A built on the GPU (cached per geometry, so repeated calls only run the SpMV)."""
from __future__ import annotations

from ._runtime import DeviceArray, EntryPoint, f32_array, f32_host, f32_scalar, i32_scalar, image_side


class forwardprojection_sparse(EntryPoint):
    def main(self, angles, rhozero, deltarho, numrhos, image) -> DeviceArray:
        a = f32_host(angles, 0)
        rz = f32_scalar(rhozero, 1)
        dr = f32_scalar(deltarho, 2)
        R = i32_scalar(numrhos, 3)
        img = f32_array(image, 4, self.torch_device)
        size = image_side(img.numel())
        g = self.ctx.geometry(a, rz, dr, R, size)
        return DeviceArray(g.sparse("fp").spmv(img))
