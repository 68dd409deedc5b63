"""CPU restatement of tomograph/synkrotomo's projector and SIRT (the oracle).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg, always as the checker or the timed CPU column --
never as a fallback for the GPU path.  The arithmetic lives in ``oracle.c`` (built by
``make -C oracle``); this module is a numpy/ctypes veneer over it.

Reference functions restated (file:line in tomograph/synkrotomo):
  preprocess       futhark/sirtlib.fut:28-44
  fp               futhark/forwardprojection.fut:21-84
  bp               futhark/backprojection.fut:20-93
  sirt             futhark/SIRT.fut:9-49
  sirt3d           futhark/SIRT3D.fut:18-28 (argument defect fixed: rho0 = rhos[0],
                   drho = rhos[1]-rhos[0], numrhos = len(rhos))
  sanity_check     futhark/tools/sanityCheck.fut:1-4
  diff             futhark/tools/diff.fut:1-7
  mse              futhark/tools/mse.fut:1-3
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double
    sig = {
        "orc_preprocess": (c_int, [_f32p, c_int, _f32p, _f32p, _f32p, _i32p]),
        "orc_fp": (None, [_f32p, c_int, c_float, c_float, c_int, _f32p, c_int, _f32p, c_int]),
        "orc_bp": (None, [_f32p, c_int, c_float, c_float, c_int, _f32p, c_int, _f32p, c_int]),
        "orc_sirt": (None, [_f32p, c_int, c_float, c_float, _f32p, c_int, _f32p, c_int, c_int, c_int]),
        "orc_sirt3d": (None, [_f32p, c_int, _f32p, c_int, _f32p, c_int, c_int, _f32p, c_int]),
        "orc_fp_f64": (None, [_f32p, c_int, c_double, c_double, c_int, _f64p, c_int, _f64p, c_int]),
        "orc_bp_f64": (None, [_f32p, c_int, c_double, c_double, c_int, _f64p, c_int, _f64p, c_int]),
        "orc_sirt_f64": (None, [_f32p, c_int, c_double, c_double, _f64p, c_int, _f64p, c_int, c_int, c_int]),
        "orc_fp_ray_line": (c_float, [c_float, c_float, c_float, c_int, c_float, c_float, c_int, _f32p]),
        "orc_bp_pixel_line": (c_float, [c_float, c_float, c_float, c_float, c_float, c_float, c_float,
                                        c_int, c_int, _f32p, _i32p, _f32p]),
        "orc_fp_csr": (ctypes.c_int64, [_f32p, c_int, c_float, c_float, c_int, c_int, _i64p, _i32p, _f32p]),
        "orc_bp_csr": (ctypes.c_int64, [_f32p, c_int, c_float, c_float, c_int, c_int, _i64p, _i32p, _f32p]),
        "orc_fp_adjoint": (None, [_f32p, c_int, c_float, c_float, c_int, c_int, _f32p, _f64p]),
        "orc_bp_oob_nonzero": (ctypes.c_int64, [_f32p, c_int, c_float, c_float, c_int, c_int]),
        "orc_sanity_check": (c_int, [_f32p, _f32p, ctypes.c_int64, c_float]),
        "orc_diff": (ctypes.c_int64, [_f32p, _f32p, ctypes.c_int64, c_float]),
        "orc_mse": (c_float, [_f32p, _f32p, ctypes.c_int64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a: np.ndarray, ptype):
    return a.ctypes.data_as(ptype)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def default_threads() -> int:
    return max(1, len(os.sched_getaffinity(0)))


def preprocess(angles):
    """sirtlib.fut:28-44 -> (cost, sint, lbase, orig, n_steep) in steep-then-flat order."""
    a = _f32(angles)
    A = a.size
    c, s, l = (np.empty(A, np.float32) for _ in range(3))
    o = np.empty(A, np.int32)
    ns = _load().orc_preprocess(_p(a, _f32p), A, _p(c, _f32p), _p(s, _f32p), _p(l, _f32p), _p(o, _i32p))
    return c, s, l, o, int(ns)


def fp(angles, rhozero, deltarho, numrhos, image, nthreads=1):
    """forwardprojection.fut:71-84 main -> [A*numrhos] float32."""
    a, img = _f32(angles), _f32(image)
    out = np.empty(a.size * int(numrhos), np.float32)
    _load().orc_fp(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(numrhos),
                   _p(img, _f32p), img.size, _p(out, _f32p), int(nthreads))
    return out


def bp(angles, rhozero, deltarho, size, projections, nthreads=1):
    """backprojection.fut:82-93 main -> [size*size] float32."""
    a, pr = _f32(angles), _f32(projections)
    out = np.empty(int(size) * int(size), np.float32)
    _load().orc_bp(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(size),
                   _p(pr, _f32p), pr.size, _p(out, _f32p), int(nthreads))
    return out


def sirt(angles, rhozero, deltarho, image, projections, iterations, nthreads=1):
    """SIRT.fut:43-49 main -> [n] float32 (the input image is not modified)."""
    a, pr = _f32(angles), _f32(projections)
    img = _f32(image).copy()
    _load().orc_sirt(_p(a, _f32p), a.size, float(rhozero), float(deltarho), _p(img, _f32p), img.size,
                     _p(pr, _f32p), pr.size, int(iterations), int(nthreads))
    return img


def sirt3d(angles, rhos, projections, iterations, size, nthreads=1):
    """SIRT3D.fut:18-28 (fixed) -> [size*size*size] float32."""
    a, rh, pr = _f32(angles), _f32(rhos), _f32(projections)
    out = np.empty(int(size) ** 3, np.float32)
    _load().orc_sirt3d(_p(a, _f32p), a.size, _p(rh, _f32p), rh.size, _p(pr, _f32p), int(iterations),
                       int(size), _p(out, _f32p), int(nthreads))
    return out


def fp_f64(angles, rhozero, deltarho, numrhos, image, nthreads=1):
    a, img = _f32(angles), _f64(image)
    out = np.empty(a.size * int(numrhos), np.float64)
    _load().orc_fp_f64(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(numrhos),
                       _p(img, _f64p), img.size, _p(out, _f64p), int(nthreads))
    return out


def bp_f64(angles, rhozero, deltarho, size, projections, nthreads=1):
    a, pr = _f32(angles), _f64(projections)
    out = np.empty(int(size) * int(size), np.float64)
    _load().orc_bp_f64(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(size),
                       _p(pr, _f64p), pr.size, _p(out, _f64p), int(nthreads))
    return out


def sirt_f64(angles, rhozero, deltarho, image, projections, iterations, nthreads=1):
    a, pr = _f32(angles), _f64(projections)
    img = _f64(image).copy()
    _load().orc_sirt_f64(_p(a, _f32p), a.size, float(rhozero), float(deltarho), _p(img, _f64p),
                         img.size, _p(pr, _f64p), pr.size, int(iterations), int(nthreads))
    return img


def fp_ray_line(cost, sint, lbase, r, rhozero, deltarho, halfsize, frame_image):
    img = _f32(frame_image)
    return float(_load().orc_fp_ray_line(float(cost), float(sint), float(lbase), int(r), float(rhozero),
                                         float(deltarho), int(halfsize), _p(img, _f32p)))


def bp_pixel_line(cost, sint, lbase, xmin, ymin, rhozero, deltarho, ncand=2, R=1 << 30, proj_row=None):
    """backprojection.fut:48-65 for one (pixel, line): (candidate bins, lengths, inner sum)."""
    s = np.empty(ncand, np.int32)
    l = np.empty(ncand, np.float32)
    row = None if proj_row is None else _f32(proj_row)
    v = _load().orc_bp_pixel_line(float(cost), float(sint), float(lbase), float(xmin), float(ymin),
                                  float(rhozero), float(deltarho), int(ncand), int(R),
                                  None if row is None else _p(row, _f32p), _p(s, _i32p), _p(l, _f32p))
    return s, l, float(v)


def fp_csr(angles, rhozero, deltarho, R, N):
    """CSR of the forward operator A (rows = rays angle-major, cols = pixels)."""
    a = _f32(angles)
    lib = _load()
    nnz = lib.orc_fp_csr(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(R), int(N), None, None, None)
    rowptr = np.empty(a.size * R + 1, np.int64)
    cols = np.empty(max(nnz, 1), np.int32)
    vals = np.empty(max(nnz, 1), np.float32)
    lib.orc_fp_csr(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(R), int(N),
                   _p(rowptr, _i64p), _p(cols, _i32p), _p(vals, _f32p))
    return rowptr, cols[:nnz], vals[:nnz]


def bp_csr(angles, rhozero, deltarho, R, N):
    """CSR of the backprojection operator B (rows = pixels, cols = sinogram entries)."""
    a = _f32(angles)
    lib = _load()
    nnz = lib.orc_bp_csr(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(R), int(N), None, None, None)
    rowptr = np.empty(N * N + 1, np.int64)
    cols = np.empty(max(nnz, 1), np.int32)
    vals = np.empty(max(nnz, 1), np.float32)
    lib.orc_bp_csr(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(R), int(N),
                   _p(rowptr, _i64p), _p(cols, _i32p), _p(vals, _f32p))
    return rowptr, cols[:nnz], vals[:nnz]


def fp_adjoint(angles, rhozero, deltarho, R, N, sino):
    """Exact A^T sino (float64 accumulation over A's entries)."""
    a, s = _f32(angles), _f32(sino)
    out = np.empty(N * N, np.float64)
    _load().orc_fp_adjoint(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(R), int(N),
                           _p(s, _f32p), _p(out, _f64p))
    return out


def bp_oob_nonzero(angles, rhozero, deltarho, R, N):
    a = _f32(angles)
    return int(_load().orc_bp_oob_nonzero(_p(a, _f32p), a.size, float(rhozero), float(deltarho), int(R), int(N)))


def _pair(sane, test):
    a, b = _f32(sane), _f32(test)
    if a.size != b.size:
        raise ValueError("sizes differ")
    return a, b


def sanity_check(sane, test, tol=0.05):
    """tools/sanityCheck.fut: every |sane - test| < tol."""
    a, b = _pair(sane, test)
    return bool(_load().orc_sanity_check(_p(a, _f32p), _p(b, _f32p), a.size, tol))


def diff(sane, test, tol=0.05):
    """tools/diff.fut: count of indices i > 0 with |sane - test| > tol."""
    a, b = _pair(sane, test)
    return int(_load().orc_diff(_p(a, _f32p), _p(b, _f32p), a.size, tol))


def mse(sane, test):
    """tools/mse.fut: sequential f32 mean of the squared differences."""
    a, b = _pair(sane, test)
    return float(_load().orc_mse(_p(a, _f32p), _p(b, _f32p), a.size))
