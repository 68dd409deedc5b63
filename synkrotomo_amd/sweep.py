"""Size sweeps of the projector in the reference's result-file format, and speedups against the
reference's own result files.

The reference times forwardprojection / backprojection on a series of image sizes and keeps
the results as ``size;microseconds`` lines (output/*/fut_fp.csv, fut_bp.csv, and the ASTRA
comparison's astra_fp.csv, astra_bp.csv; a file may hold several runs of a size).  Its
datasets (futhark/data/fpinputf32rad512: 805 angles pi/4 + k*pi/1610, 725 bins of width 1,
rhozero -362, a 512x512 image) follow one rule per size N: A = ceil(pi/2 * N) angles
pi/4 + k*pi/(2A) in [pi/4, 3pi/4), R = ceil(sqrt(2) * N) bins, rhozero = -(R-1)/2, deltarho 1.
``--angles full`` spreads the same A angles over [0, pi) instead (k*pi/A).

    python -m synkrotomo_amd.sweep [--sizes 64,128,...] [--angles quarter|full] [--runs 10]
                                   [--out DIR] [--compare DIR]

For each size: one geometry, a random image in [0, 1) and its sinogram resident on the GPU;
``--runs`` timed calls of A.x and of B.r after one warm-up (HIP events on the launch stream;
the per-call transpose of the image for the flat lines is included, the one-time line table
and kernel choice are not).  Writes DIR/fp.csv and DIR/bp.csv (``size;us``, mean of the runs)
and prints one JSON line with GUPS per size; with ``--compare`` the median of each size in
DIR/fut_fp.csv, fut_bp.csv (and astra_*.csv when present) divided by this build's time.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
from collections import defaultdict

import numpy as np

REFERENCE_SIZES = (64, 128, 256, 512, 1024, 1500, 2000, 2048, 2500, 3000, 3500, 4000, 4096)


def sweep_geometry(N: int, angles: str = "quarter"):
    """(angles f32 [A], rhozero, deltarho, R) of the reference's dataset rule for size N."""
    A = int(math.ceil(math.pi / 2 * N))
    k = np.arange(A, dtype=np.float64)
    if angles == "quarter":
        th = math.pi / 4 + k * math.pi / (2 * A)
    elif angles == "full":
        th = k * math.pi / A
    else:
        raise ValueError("angles must be 'quarter' or 'full'")
    R = int(math.ceil(math.sqrt(2.0) * N))
    return th.astype(np.float32), -(R - 1) / 2.0, 1.0, R


def read_results(path: str) -> dict:
    """``size;us`` lines -> {size: [us, ...]} (blank lines and others ignored)."""
    out = defaultdict(list)
    with open(path) as f:
        for line in f:
            parts = line.strip().split(";")
            if len(parts) != 2:
                continue
            try:
                out[int(parts[0])].append(float(parts[1]))
            except ValueError:
                continue
    return dict(out)


def write_results(path: str, rows) -> None:
    with open(path, "w") as f:
        for size, us in rows:
            f.write(f"{size};{us:.2f}\n")


def speedups(ours: dict, theirs: dict) -> dict:
    """{size: median(theirs) / ours} for the sizes in both."""
    return {n: round(float(np.median(theirs[n])) / ours[n], 2) for n in sorted(ours) if n in theirs}


def run_sweep(sizes, angles: str = "quarter", runs: int = 10, ctx=None) -> dict:
    import torch
    from .autotune import tuned
    from .projector import Context
    ctx = ctx or Context.default(0)
    fp_us, bp_us, gups = {}, {}, {}
    gen = torch.Generator(device="cuda").manual_seed(0)
    for N in sizes:
        th, rz, dr, R = sweep_geometry(N, angles)
        geo = ctx.geometry(th, rz, dr, R, N)
        tuned(geo)
        x = torch.rand((N, N), device="cuda", generator=gen)
        sino = geo.fp(x)
        img = geo.bp(sino)
        res = {}
        for name, fn in (("fp", lambda: geo.fp(x, out=sino)), ("bp", lambda: geo.bp(sino, out=img))):
            fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(runs):
                fn()
            e1.record()
            torch.cuda.synchronize()
            res[name] = e0.elapsed_time(e1) * 1e3 / runs
        fp_us[N], bp_us[N] = res["fp"], res["bp"]
        upd = float(N) * N * th.size
        gups[N] = {"angles": int(th.size), "bins": R, "fp_gups": round(upd / res["fp"] / 1e3, 1),
                   "bp_gups": round(upd / res["bp"] / 1e3, 1)}
        del geo, x, sino, img
        torch.cuda.empty_cache()
    return {"fp_us": fp_us, "bp_us": bp_us, "gups": gups}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m synkrotomo_amd.sweep")
    ap.add_argument("--sizes", default=",".join(map(str, REFERENCE_SIZES)))
    ap.add_argument("--angles", choices=("quarter", "full"), default="quarter")
    ap.add_argument("--runs", type=int, default=10)
    ap.add_argument("--out", default=".", help="directory for fp.csv and bp.csv")
    ap.add_argument("--compare", help="directory with the reference's fut_*.csv / astra_*.csv")
    args = ap.parse_args(argv)
    sizes = [int(s) for s in args.sizes.split(",") if s.strip()]
    if any(n < 2 or n % 2 for n in sizes):
        raise SystemExit("sizes must be even and >= 2")
    r = run_sweep(sizes, args.angles, args.runs)
    os.makedirs(args.out, exist_ok=True)
    write_results(os.path.join(args.out, "fp.csv"), sorted(r["fp_us"].items()))
    write_results(os.path.join(args.out, "bp.csv"), sorted(r["bp_us"].items()))
    summary = {"angles": args.angles, "runs": args.runs, "gups": r["gups"]}
    if args.compare:
        cmp = {}
        for kind in ("fp", "bp"):
            for who in ("fut", "astra"):
                path = os.path.join(args.compare, f"{who}_{kind}.csv")
                if os.path.exists(path):
                    cmp[f"{who}_{kind}"] = speedups(r[f"{kind}_us"], read_results(path))
        summary["speedup_vs_reference_files"] = cmp
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
