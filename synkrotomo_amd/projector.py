"""Device-level projector API: contexts, geometries and the operators on torch tensors.

torch supplies device memory and the current stream; every operator here is one call
into libsynkro.so (hand-written HIP).  Images are float32 ``[..., N, N]`` (or flat
``[..., N*N]``), sinograms float32 ``[..., A, R]`` (or flat), angle-major.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np
import torch

from . import _native as nat


def _stream(device: int):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """A libsynkro context (``sk_ctx``) bound to one device; use one per host thread."""

    _tls = threading.local()

    def __init__(self, device: int = 0):
        L = nat.lib()
        if not torch.cuda.is_available():
            raise RuntimeError("synkrotomo_amd needs a ROCm GPU; none is visible")
        h = ctypes.c_void_p()
        nat.check(L.sk_create(int(device), ctypes.byref(h)), None, "sk_create")
        self.handle = h
        self.device = int(device)
        self._geoms: dict = {}

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        ctxs = getattr(cls._tls, "ctxs", None)
        if ctxs is None:
            ctxs = cls._tls.ctxs = {}
        c = ctxs.get(device)
        if c is None:
            c = ctxs[device] = cls(device)
        return c

    # options -------------------------------------------------------------
    def set_option(self, option: int, value: int) -> None:
        nat.check(nat.lib().sk_set_option(self.handle, option, int(value)), self.handle, "sk_set_option")

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        nat.check(nat.lib().sk_get_option(self.handle, option, ctypes.byref(v)), self.handle)
        return v.value

    @property
    def exact(self) -> bool:
        return bool(self.get_option(nat.SK_OPT_EXACT))

    @exact.setter
    def exact(self, flag: bool) -> None:
        self.set_option(nat.SK_OPT_EXACT, 1 if flag else 0)

    # geometries ----------------------------------------------------------
    def geometry(self, angles, rhozero, deltarho, numrhos, size) -> "Geometry":
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float32).ravel())
        key = (a.tobytes(), float(np.float32(rhozero)), float(np.float32(deltarho)), int(numrhos), int(size))
        g = self._geoms.get(key)
        if g is None:
            g = Geometry(self, a, rhozero, deltarho, numrhos, size)
            if len(self._geoms) > 64:
                self._geoms.clear()
            self._geoms[key] = g
        return g

    def close(self) -> None:
        for g in self._geoms.values():
            g.close()
        self._geoms.clear()
        if self.handle:
            nat.lib().sk_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Geometry:
    """Parallel-beam geometry: angles [A], bins rho = rho0 + r*drho (r < R), image N x N."""

    def __init__(self, ctx: Context, angles, rhozero, deltarho, numrhos, size):
        self.ctx = ctx
        self.angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float32).ravel())
        self.rho0 = float(np.float32(rhozero))
        self.drho = float(np.float32(deltarho))
        h = ctypes.c_void_p()
        L = nat.lib()
        nat.check(L.sk_geom_create(ctx.handle, self.angles.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   self.angles.size, self.rho0, self.drho, int(numrhos), int(size),
                                   ctypes.byref(h)), ctx.handle, "geometry")
        self.handle = h
        A, R, N, ns, nc = (ctypes.c_int() for _ in range(5))
        L.sk_geom_query(h, *(ctypes.byref(v) for v in (A, R, N, ns, nc)))
        self.A, self.R, self.N, self.n_steep, self.ncand = A.value, R.value, N.value, ns.value, nc.value

    def close(self) -> None:
        for m in getattr(self, "_mats", {}).values():
            m.close()
        self._mats = {}
        if self.handle:
            nat.lib().sk_geom_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sparse(self, kind: str) -> "Matrix":
        """The explicit matrix of this geometry ("fp": A, "bp": B, "adj": A^T), built once and
        cached."""
        mats = self.__dict__.setdefault("_mats", {})
        m = mats.get(kind)
        if m is None:
            m = mats[kind] = Matrix(self, kind)
        return m

    def drop_sparse(self) -> None:
        for m in getattr(self, "_mats", {}).values():
            m.close()
        self._mats = {}

    def sparse_nnz(self, kind: str) -> int:
        """nnz of the explicit matrix without building it."""
        which = _MAT_KINDS[kind]
        v = ctypes.c_int64()
        nat.check(nat.lib().sk_sparse_count(self.ctx.handle, self.handle, which, ctypes.byref(v),
                                            _stream(self.ctx.device)), self.ctx.handle, "sparse count")
        return v.value

    # helpers ---------------------------------------------------------------
    @property
    def device(self) -> torch.device:
        return torch.device("cuda", self.ctx.device)

    def _check(self, t: torch.Tensor, per: int, name: str) -> int:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if t.dtype != torch.float32:
            raise TypeError(f"{name} must be float32 (got {t.dtype})")
        if t.device != self.device:
            raise ValueError(f"{name} must be on {self.device} (got {t.device})")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if t.numel() == 0 or t.numel() % per:
            raise ValueError(f"{name} has {t.numel()} elements, not a multiple of {per}")
        return t.numel() // per

    def _img_shape(self, sino: torch.Tensor):
        lead = sino.shape[:-2] if sino.dim() >= 2 and tuple(sino.shape[-2:]) == (self.A, self.R) else \
            (sino.numel() // (self.A * self.R),)
        if sino.dim() == 1:
            return (self.N * self.N,)
        return (*lead, self.N, self.N)

    def _sino_shape(self, img: torch.Tensor):
        if img.dim() == 1:
            return (self.A * self.R,)
        lead = img.shape[:-2] if tuple(img.shape[-2:]) == (self.N, self.N) else (img.numel() // (self.N * self.N),)
        return (*lead, self.A, self.R)

    # operators -------------------------------------------------------------
    def fp(self, img: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """A.x (forwardprojection.main)."""
        batch = self._check(img, self.N * self.N, "image")
        if out is None:
            out = torch.empty(self._sino_shape(img), dtype=torch.float32, device=self.device)
        elif self._check(out, self.A * self.R, "out") != batch:
            raise ValueError("out batch mismatch")
        nat.check(nat.lib().sk_fp(self.ctx.handle, self.handle, batch, _ptr(img), _ptr(out),
                                  _stream(self.ctx.device)), self.ctx.handle, "forward projection")
        return out

    def bp(self, sino: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """B.r (backprojection.main)."""
        batch = self._check(sino, self.A * self.R, "projections")
        if out is None:
            out = torch.empty(self._img_shape(sino), dtype=torch.float32, device=self.device)
        elif self._check(out, self.N * self.N, "out") != batch:
            raise ValueError("out batch mismatch")
        nat.check(nat.lib().sk_bp(self.ctx.handle, self.handle, batch, _ptr(sino), _ptr(out),
                                  _stream(self.ctx.device)), self.ctx.handle, "backprojection")
        return out

    def bp_adjoint(self, sino: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """A^T.r, the exact adjoint of fp."""
        batch = self._check(sino, self.A * self.R, "projections")
        if out is None:
            out = torch.empty(self._img_shape(sino), dtype=torch.float32, device=self.device)
        elif self._check(out, self.N * self.N, "out") != batch:
            raise ValueError("out batch mismatch")
        nat.check(nat.lib().sk_bp_adjoint(self.ctx.handle, self.handle, batch, _ptr(sino), _ptr(out),
                                          _stream(self.ctx.device)), self.ctx.handle, "adjoint")
        return out

    def bp_transpose(self, img: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """B^T.v, the exact adjoint of bp."""
        batch = self._check(img, self.N * self.N, "image")
        if out is None:
            out = torch.empty(self._sino_shape(img), dtype=torch.float32, device=self.device)
        elif self._check(out, self.A * self.R, "out") != batch:
            raise ValueError("out batch mismatch")
        nat.check(nat.lib().sk_bp_transpose(self.ctx.handle, self.handle, batch, _ptr(img), _ptr(out),
                                            _stream(self.ctx.device)), self.ctx.handle, "bp transpose")
        return out

    def sirt_(self, img: torch.Tensor, sino: torch.Tensor, iterations: int) -> torch.Tensor:
        """SIRT in place on img (SIRT.main; batch > 1 = independent slices)."""
        batch = self._check(img, self.N * self.N, "image")
        if self._check(sino, self.A * self.R, "projections") != batch:
            raise ValueError("image and projections batch sizes differ")
        if int(iterations) < 0:
            raise ValueError("iterations must be >= 0")
        nat.check(nat.lib().sk_sirt(self.ctx.handle, self.handle, batch, _ptr(img), _ptr(sino),
                                    int(iterations), _stream(self.ctx.device)), self.ctx.handle, "SIRT")
        return img

    def autotune(self, batch: int = 1):
        """Time the kernel variants on this geometry on calls of `batch` images, keep the
        fastest; returns (fp, bp)."""
        fv, bv = ctypes.c_int(), ctypes.c_int()
        nat.check(nat.lib().sk_autotune_batch(self.ctx.handle, self.handle, int(batch), ctypes.byref(fv),
                                              ctypes.byref(bv), _stream(self.ctx.device)),
                  self.ctx.handle, "autotune")
        return fv.value, bv.value

    def set_variants(self, fp_variant: int = -1, bp_variant: int = -1) -> None:
        nat.check(nat.lib().sk_geom_set_variants(self.handle, int(fp_variant), int(bp_variant)),
                  self.ctx.handle, "set variants")

    def sirt_weights(self):
        """(1/rowsum(A) as [A, R], 1/colsum(B) as [N, N]), 0 where the sum is 0."""
        rw = torch.empty((self.A, self.R), dtype=torch.float32, device=self.device)
        cw = torch.empty((self.N, self.N), dtype=torch.float32, device=self.device)
        nat.check(nat.lib().sk_sirt_weights(self.ctx.handle, self.handle, _ptr(rw), _ptr(cw),
                                            _stream(self.ctx.device)), self.ctx.handle, "SIRT weights")
        return rw, cw

    def sirt_step_(self, img, sino, rowinv, colinv, q=None) -> torch.Tensor:
        """One SIRT iteration in place on img with the given weights."""
        batch = self._check(img, self.N * self.N, "image")
        if self._check(sino, self.A * self.R, "projections") != batch:
            raise ValueError("image and projections batch sizes differ")
        self._check(rowinv, self.A * self.R, "rowinv")
        self._check(colinv, self.N * self.N, "colinv")
        if q is None:
            q = torch.empty((batch, self.A, self.R), dtype=torch.float32, device=self.device)
        nat.check(nat.lib().sk_sirt_step(self.ctx.handle, self.handle, batch, _ptr(img), _ptr(sino),
                                         _ptr(rowinv), _ptr(colinv), _ptr(q), _stream(self.ctx.device)),
                  self.ctx.handle, "SIRT step")
        return img


_MAT_KINDS = {"fp": nat.SK_MAT_FP, "bp": nat.SK_MAT_BP, "adj": nat.SK_MAT_ADJ}


class Matrix:
    """An explicit system matrix in CSR form, built on the GPU (``sk_sparse_build``).

    kind "fp": A, rows = rays (angle*R + bin), columns = pixels.
    kind "bp": B, rows = pixels, columns = sinogram entries (the reference backprojector).
    kind "adj": A^T (A in CSC form), rows = pixels, columns = rays; A's entries bit for bit.
    """

    def __init__(self, geom: "Geometry", kind: str):
        which = _MAT_KINDS[kind]
        self.geom = geom
        self.kind = kind
        h = ctypes.c_void_p()
        nat.check(nat.lib().sk_sparse_build(geom.ctx.handle, geom.handle, which, ctypes.byref(h),
                                            _stream(geom.ctx.device)), geom.ctx.handle, "sparse build")
        self.handle = h
        rows, cols, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        nat.lib().sk_mat_info(h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz))
        self.rows, self.cols, self.nnz = rows.value, cols.value, nnz.value

    def close(self) -> None:
        if self.handle:
            nat.lib().sk_mat_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csr(self):
        """(rowptr int64 [rows+1], cols int32 [nnz], vals float32 [nnz]) as new device tensors."""
        dev = self.geom.device
        rp = torch.empty(self.rows + 1, dtype=torch.int64, device=dev)
        c = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=dev)
        v = torch.empty(max(self.nnz, 1), dtype=torch.float32, device=dev)
        nat.check(nat.lib().sk_mat_copy(self.geom.ctx.handle, self.handle, _ptr(rp), _ptr(c), _ptr(v),
                                        _stream(self.geom.ctx.device)), self.geom.ctx.handle, "matrix copy")
        return rp, c[: self.nnz], v[: self.nnz]

    @property
    def spmv_variant(self) -> int:
        """The SpMV kernel in use (-1 until the first spmv call times the candidates)."""
        return int(nat.lib().sk_mat_spmv_variant(self.handle))

    def set_spmv_variant(self, variant: int) -> None:
        """0: wave per row, 1: lane per row (lockstep), 2: wave per row with 16-byte streams,
        3 / 4: as 0 / 2 with the steep rays' gathers from x^T (A only), 7: lane per row on a
        SELL-64 copy of the matrix (built on first use, about the CSR's size), 8: as 7 with the
        flat rays' gathers from x^T (A only), 9: as 7 with 16-bit column steps (6 bytes per
        entry; ValueError on use when a row's consecutive columns are more than 32767 apart);
        10: row slots, a copy whose 64-row slices walk in step (A: frame rows, B: lines), so a
        wave's gathers fall on one frame row or one line (A and B only; B needs deltarho in
        [0.71, 1.41)); -1: time the candidates on next use and keep the fastest."""
        nat.check(nat.lib().sk_mat_set_spmv_variant(self.handle, int(variant)), self.geom.ctx.handle)

    def spmv(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        g = self.geom
        batch = g._check(x, self.cols, "input")
        if out is None:
            out = torch.empty((batch, self.rows) if batch > 1 else (self.rows,), dtype=torch.float32,
                              device=g.device)
        elif g._check(out, self.rows, "out") != batch:
            raise ValueError("out batch mismatch")
        nat.check(nat.lib().sk_spmv(g.ctx.handle, self.handle, batch, _ptr(x), _ptr(out),
                                    _stream(g.ctx.device)), g.ctx.handle, "SpMV")
        return out


def sirt_batched(groups, iterations: int, ctx: Context | None = None) -> None:
    """Advance every problem ``iterations`` SIRT iterations in one batched launch sequence.

    groups: iterable of (geometry, images [b, N, N], sinograms [b, A, R]) -- each group any
    number of independent problems of one geometry; images are updated in place."""
    groups = list(groups)
    if not groups:
        raise ValueError("no problem groups")
    ctx = ctx or groups[0][0].ctx
    arr = (nat.SkProblem * len(groups))()
    keep = []
    for k, (geom, img, sino) in enumerate(groups):
        if geom.ctx is not ctx:
            raise ValueError("all geometries must belong to one context")
        b = geom._check(img, geom.N * geom.N, "image")
        if geom._check(sino, geom.A * geom.R, "projections") != b:
            raise ValueError("image and projections batch sizes differ")
        arr[k].geom = geom.handle.value
        arr[k].batch = b
        arr[k].d_img = img.data_ptr()
        arr[k].d_sino = sino.data_ptr()
        keep.append((img, sino))
    nat.check(nat.lib().sk_sirt_batched(ctx.handle, ctypes.cast(arr, ctypes.c_void_p), len(groups),
                                        int(iterations), _stream(ctx.device)), ctx.handle, "batched SIRT")


def uniform_angles(A: int) -> np.ndarray:
    """theta_k = k*pi/A as float32, k < A (angles uniform in [0, pi))."""
    return (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)


def centred_bins(N: int):
    """(R, rho0) with R = ceil(sqrt(2)*N) bins of width 1 centred on the image."""
    R = int(np.ceil(np.sqrt(2.0) * N))
    return R, -(R - 1) / 2.0
