"""Regenerate the committed golden fixtures of tests/golden/.

    python tests/golden/make_golden.py

Writes
  fpinput512.npz   the reference's own benchmark input futhark/data/fpinputf32rad512
                   (805 angles in [pi/4, 3pi/4), rhozero -362, deltarho 1, 725 bins, a
                   512x512 image of values k/255), stored as angles + uint8 k, plus the
                   oracle's fp and bp outputs on it (every 8th angle / image row, and
                   SHA-256 of the full float32 outputs).
  c1.npz           the 64x64 / 90-angle / 91-bin configuration (10-ellipse phantom, seed
                   0): angles, phantom, and the oracle's fp, bp and 10-iteration SIRT.
  ref_results_rtx2080ti/{fut,astra}_{fp,bp}.csv
                   the reference's own timing results (output/20190523_065504: Futhark
                   and ASTRA on an RTX 2080 Ti, sizes 64..4096, 10 runs each, size;us),
                   copied unchanged for synkrotomo_amd.sweep --compare.
The expected outputs come from the oracle (oracle/oracle.c), the CPU restatement of
the reference; the reference itself (Futhark) cannot be built here.  Reading the
reference's input file needs /root/reference; the outputs are committed so that nothing
at test time needs it.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from synkrotomo_amd.phantom import ellipse_phantom  # noqa: E402

REF_FIXTURE = "/root/reference/futhark/data/fpinputf32rad512"
REF_RESULTS = "/root/reference/output/20190523_065504"


def parse_futhark_values(text: str):
    """Futhark text value format: whitespace-separated values, arrays as [v, v, ...],
    scalars with f32/i32 suffixes."""
    vals = []
    i, n = 0, len(text)
    while i < n:
        ch = text[i]
        if ch.isspace():
            i += 1
            continue
        if ch == "[":
            j = text.index("]", i)
            body = text[i + 1:j]
            items = [t.strip() for t in body.split(",") if t.strip()]
            vals.append(items)
            i = j + 1
        else:
            j = i
            while j < n and not text[j].isspace():
                j += 1
            vals.append(text[i:j])
            i = j
    return vals


def lit(s: str):
    if s.endswith("f32"):
        return np.float32(float(s[:-3]))
    if s.endswith("i32"):
        return int(s[:-3])
    raise ValueError(s)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def make_fixture512():
    vals = parse_futhark_values(open(REF_FIXTURE).read())
    angles = np.array([lit(s) for s in vals[0]], np.float32)
    rhozero, deltarho, numrhos = lit(vals[1]), lit(vals[2]), lit(vals[3])
    image = np.array([lit(s) for s in vals[4]], np.float32)
    k = np.rint(image.astype(np.float64) * 255.0).astype(np.uint8)
    assert np.array_equal((k.astype(np.float64) / 255.0).astype(np.float32), image), "image is not k/255"
    nt = oracle.default_threads()
    fp = oracle.fp(angles, rhozero, deltarho, numrhos, image, nthreads=nt)
    bp = oracle.bp(angles, rhozero, deltarho, 512, fp, nthreads=nt)
    np.savez_compressed(
        os.path.join(HERE, "fpinput512.npz"),
        angles=angles, rhozero=np.float32(rhozero), deltarho=np.float32(deltarho),
        numrhos=np.int32(numrhos), image_k=k.reshape(512, 512),
        fp_rows=fp.reshape(angles.size, numrhos)[::8], fp_sha=sha(fp),
        bp_rows=bp.reshape(512, 512)[::8], bp_sha=sha(bp))
    print("fpinput512: angles", angles.size, "fp sha", sha(fp)[:12], "bp sha", sha(bp)[:12])


def make_c1():
    N, A, R, rz = 64, 90, 91, -45.0
    angles = (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)
    phantom = ellipse_phantom(N, 10, seed=0).ravel()
    fp = oracle.fp(angles, rz, 1.0, R, phantom)
    bp = oracle.bp(angles, rz, 1.0, N, fp)
    sirt10 = oracle.sirt(angles, rz, 1.0, np.zeros(N * N, np.float32), fp, 10)
    np.savez_compressed(os.path.join(HERE, "c1.npz"), angles=angles, phantom=phantom, fp=fp, bp=bp,
                        sirt10=sirt10, rhozero=np.float32(rz), deltarho=np.float32(1.0),
                        numrhos=np.int32(R), size=np.int32(N))
    print("c1: fp sha", sha(fp)[:12], "sirt10 sha", sha(sirt10)[:12])


def copy_ref_results():
    out = os.path.join(HERE, "ref_results_rtx2080ti")
    os.makedirs(out, exist_ok=True)
    for name in ("fut_fp.csv", "fut_bp.csv", "astra_fp.csv", "astra_bp.csv"):
        with open(os.path.join(REF_RESULTS, name)) as f, open(os.path.join(out, name), "w") as g:
            g.write(f.read())


if __name__ == "__main__":
    oracle.build()
    make_c1()
    make_fixture512()
    copy_ref_results()
