#!/usr/bin/env python3
"""Benchmark: dense A.x and A^T.r (B.r) at 2048^2, 2048 angles, 2897 bins on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one forward projection plus one backprojection of a 2048x2048 float32
50-ellipse phantom (seeded) over 2048 angles uniform in [0, pi) and 2897 unit bins
centred on the image (rhozero = -1448): the configuration the headline metric is quoted
on.  Inputs are resident in HBM before timing.  Prints ONE JSON line.

value      = GUPS = 2 * N^2 * A * K / (wall time of K steps) / 1e9   (fwd and back updates)
roofline   = VALU issue of the dominant kernel: its VALU wave-instructions per launch
             (SQ_INSTS_VALU, profiles/counters.json) / its mean launch time (HIP events on the
             launch stream), against 2 per clock per CU (the published FP32 vector rate); the
             HBM bytes it moves (FETCH_SIZE + WRITE_SIZE) and their fraction of 8 TB/s beside it
cpu_baseline = the oracle (C restatement of the reference, OpenMP) on a bounded sample
             of the same workload on this host's cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, A, R, RHO0 = 2048, 2048, 2897, -1448.0
NNZ_C2 = 10_937_047_089          # exact nnz(A) at this geometry (oracle CSR count)
HBM_PEAK_GBS = 8000.0            # MI355X published HBM3E bandwidth
# VALU issue peak: the published 157.3 TFLOP/s FP32 vector rate is 256 CUs x 2.4 GHz x 2 wave64
# VALU instructions per clock per CU (gfx950 issues two VALU instructions per SIMD quad-cycle,
# counter SQ_ACTIVE_INST_VALU2) x 64 lanes x 2 FLOP (FMA); in wave64 VALU instructions per second:
VALU_PEAK_GINST = 157.3e12 / (64 * 2) / 1e9   # 1228.9
PHANTOM_SEED, PHANTOM_ELLIPSES = 2048, 50


def alg_bytes() -> float:
    return 4.0 * NNZ_C2 + 4.0 * (N * N + A * R)


def roofline_of(dom: str, ms: dict, exact: bool, variants) -> dict:
    """VALU-issue roofline of the dominant kernel (and the other one beside it): its VALU wave
    instructions per launch (SQ_INSTS_VALU) and HBM bytes per launch (FETCH_SIZE and WRITE_SIZE,
    calibrated on a copy kernel), both read from profiles/counters.json (the rocprofv3 passes
    over this benchmark's own steps, kernels and variants named there), divided by the launch
    time measured live here with HIP events on the launch stream.  The counts are properties of
    the kernel and the geometry (no data-dependent control flow), so they hold across runs of
    the same kernel variant; a run on other variants or in exact mode reports no counters."""
    path = os.path.join(ROOT, "profiles", "counters.json")
    cnt = {}
    if os.path.exists(path):
        with open(path) as f:
            cnt = json.load(f)
    out = {}
    for name in ("fp", "bp"):
        k = cnt.get(name, {})
        ok = (not exact and variants is not None and k.get("variant") == variants[0 if name == "fp" else 1])
        t = ms[name] * 1e-3
        e = {"kernel": k.get("kernel") if ok else None, "ms": round(ms[name], 4)}
        if ok:
            valu = float(k["valu_insts_per_launch"])
            hbm = float(k["hbm_bytes_per_launch"])
            e.update({"valu_insts_per_launch": valu,
                      "valu_ginst_s": round(valu / t / 1e9, 1),
                      "valu_frac": round(valu / t / 1e9 / VALU_PEAK_GINST, 4),
                      "hbm_bytes_per_launch": hbm,
                      "hbm_GBs": round(hbm / t / 1e9, 1),
                      "hbm_frac": round(hbm / t / 1e9 / HBM_PEAK_GBS, 4),
                      "alg_bytes_GBs": round(alg_bytes() / t / 1e9, 1)})
        out[name] = e
    d = out[dom]
    return {"bound": "valu", "kernel": dom,
            "achieved": d.get("valu_ginst_s"), "peak": round(VALU_PEAK_GINST, 1), "unit": "Gwave-inst/s",
            "frac": d.get("valu_frac"), "traffic": d.get("hbm_bytes_per_launch"),
            "hbm_frac": d.get("hbm_frac"), "alg_bytes_per_launch": alg_bytes(),
            "peak_source": "157.3 TFLOP/s FP32 vector / (64 lanes x 2 FLOP per FMA): 2 VALU "
                           "wave-instructions per clock per CU at 2.4 GHz on 256 CUs",
            "counters": "profiles/counters.json", "fp": out["fp"], "bp": out["bp"]}


def host_threads() -> tuple[int, int]:
    """(threads used, cores in this process's affinity mask).  The oracle runs on every core of
    the mask unless OMP_NUM_THREADS names the host's CPU share (the GPU hosts here give each GPU
    a share of 16 cores and set it)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (int(env) if env.isdigit() and int(env) > 0 else aff), aff


def _median_time(fn, runs: int):
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(angles, phantom, sino, runs: int = 3) -> dict:
    """The oracle (C restatement of the reference, OpenMP over rays / pixels) on the host:
    fp over all 2048 angles; bp over every 4th angle (512 of 2048; its cost is linear in the
    angle count) scaled to 2048; SIRT per iteration from the difference of a 3- and a 1-iteration
    run on every 8th angle (256 angles) scaled to 2048.  Median of `runs` for fp and bp."""
    import oracle
    nt, aff = host_threads()
    say = lambda m: print(f"[bench] cpu baseline: {m}", file=sys.stderr, flush=True)
    say(f"{nt} threads (affinity {aff}), fp 2048 angles x{runs}")
    fp_s, fp_all = _median_time(lambda: oracle.fp(angles, RHO0, 1.0, R, phantom, nthreads=nt), runs)
    sub = np.arange(0, A, 4)
    ang4, sino4 = angles[sub], sino.reshape(A, R)[sub].ravel()
    say(f"bp 512 angles x{runs}")
    bp4_s, bp_all = _median_time(lambda: oracle.bp(ang4, RHO0, 1.0, N, sino4, nthreads=nt), runs)
    bp_s = bp4_s * A / ang4.size
    sub8 = np.arange(0, A, 8)
    ang8, sino8 = angles[sub8], sino.reshape(A, R)[sub8].ravel()
    zero = np.zeros(N * N, np.float32)
    say("SIRT 1 and 3 iterations on 256 angles")
    t0 = time.perf_counter()
    oracle.sirt(ang8, RHO0, 1.0, zero, sino8, 1, nthreads=nt)
    t1 = time.perf_counter()
    oracle.sirt(ang8, RHO0, 1.0, zero, sino8, 3, nthreads=nt)
    t2 = time.perf_counter()
    it8 = ((t2 - t1) - (t1 - t0)) / 2.0
    upd = float(N) * N * A
    return {"value": round(2.0 * upd / (fp_s + bp_s) / 1e9, 3), "unit": "GUPS", "cores": nt,
            "affinity_cores": aff, "kind": "port", "runs": runs,
            "sample": (f"oracle at 2048^2, 2897 bins, {nt} threads: fp over all 2048 angles (median of "
                       f"{runs}); bp over every 4th angle, scaled x4 (median of {runs}); SIRT per "
                       "iteration from 3- minus 1-iteration runs over every 8th angle, scaled x8"),
            "fp_s": round(fp_s, 3), "bp_s": round(bp_s, 3),
            "fp_runs_s": [round(t, 3) for t in fp_all], "bp_512_runs_s": [round(t, 3) for t in bp_all],
            "fp_gups": round(upd / fp_s / 1e9, 3), "bp_gups": round(upd / bp_s / 1e9, 3),
            "sirt_ms_per_iter": round(it8 * 1e3 * A / ang8.size, 1),
            "sirt_ms_per_iter_256_angles": round(it8 * 1e3, 1)}


def sparse_leg(geo, x, sino, K: int) -> dict:
    """Explicit CSR A and B at the same geometry, built and timed one at a time (each is
    ~87 GB at 2048^2).  SpMV bytes = 8*nnz + 8*(rows+1) + 4*(rows+cols)."""
    import torch
    res = {}
    for kind, inp in (("fp", x.reshape(-1)), ("bp", sino.reshape(-1))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m = geo.sparse(kind)
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        out = torch.empty(m.rows, device="cuda")
        t0 = time.perf_counter()
        m.spmv(inp, out=out)  # first call: times the kernels (building the SELL-64 copy) and keeps one
        torch.cuda.synchronize()
        first_ms = (time.perf_counter() - t0) * 1e3
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(K):
            m.spmv(inp, out=out)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / K
        nbytes = 8.0 * m.nnz + 8.0 * (m.rows + 1) + 4.0 * (m.rows + m.cols)
        res[kind] = {"ms": round(ms, 3), "gups": round(float(N) * N * A / (ms * 1e-3) / 1e9, 2),
                     "build_ms": round(build_ms, 1), "first_call_ms": round(first_ms, 1),
                     "variant": m.spmv_variant, "nnz": m.nnz,
                     "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        geo.drop_sparse()
        del out
        torch.cuda.empty_cache()
    return res


def sirt_leg(geo, x_true, sino, iters: int = 50) -> dict:
    """Plain SIRT at the same geometry: ``iters`` iterations from a zero image on the
    phantom's sinogram (SIRT.main), timed with HIP events on the launch stream after one
    untimed call (weights and graph capture).  Reports the data residual ||Ax - b|| / ||b||
    and the image error against the phantom, both in FP64."""
    import torch
    img = torch.zeros((geo.N, geo.N), device="cuda")
    geo.sirt_(img, sino, iters)  # SIRT weights, graph capture, warm caches
    img.zero_()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    geo.sirt_(img, sino, iters)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    res = float((geo.fp(img) - sino).double().norm() / sino.double().norm())
    err = float((img - x_true).double().norm() / x_true.double().norm())
    return {"iterations": iters, "ms": round(ms, 3), "ms_per_iter": round(ms / iters, 3),
            "gups": round(2.0 * float(N) * N * A * iters / (ms * 1e-3) / 1e9, 2),
            "residual": round(res, 6), "image_rel_err": round(err, 6)}


def stream3d_leg(ctx, S: int = 512, N: int = 512, A: int = 512, iters: int = 10, chunk: int = 64) -> dict:
    """Out-of-core SIRT3D: S slices of N^2 (A angles, centred unit bins) from host memory in
    chunks of ``chunk`` slices over three streams, against the same SIRT with the stack
    resident (compute only) and against upload + SIRT + download in series."""
    import numpy as np
    import torch
    from synkrotomo_amd.autotune import tuned
    from synkrotomo_amd.projector import centred_bins, uniform_angles
    from synkrotomo_amd.stream3d import sirt3d_streamed
    R, rz = centred_bins(N)
    geo = ctx.geometry(uniform_angles(A), rz, 1.0, R, N)
    tuned(geo, batch=chunk)  # the calls run chunk (and S) slices at a time
    proj = np.random.default_rng(5).random((S, A, R), dtype=np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d = torch.from_numpy(proj).cuda()
    x = torch.zeros(S, N, N, device="cuda")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    geo.sirt_(x, d, iters)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ref = x.cpu().numpy()
    t3 = time.perf_counter()
    x.zero_()
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    geo.sirt_(x, d, iters)                      # warm: the compute-only time
    torch.cuda.synchronize()
    compute = time.perf_counter() - t4
    del d, x
    torch.cuda.empty_cache()
    sirt3d_streamed(geo, proj[: 2 * chunk], iters, chunk)   # graphs for both buffer sets
    t5 = time.perf_counter()
    out = sirt3d_streamed(geo, proj, iters, chunk)
    streamed = time.perf_counter() - t5
    return {"slices": S, "N": N, "angles": A, "bins": R, "iterations": iters, "chunk": chunk,
            "streamed_s": round(streamed, 4), "resident_compute_s": round(compute, 4),
            "serial_upload_sirt_download_s": round((t1 - t0) + (t2 - t1) + (t3 - t2), 4),
            "bitwise_equal_to_resident": bool(np.array_equal(out, ref))}


def normalize_leg(ctx, A: int = 2048, R: int = 2896, batch: int = 16, reps: int = 20) -> dict:
    """Flat/dark normalisation and -log of a [batch, A, R] count stack (16 frames each of flat
    and dark): HBM-bound, 8 bytes per entry (read raw, write out)."""
    import torch
    from synkrotomo_amd.preprocess import normalize
    g = torch.Generator(device="cuda").manual_seed(1)
    raw = torch.rand((batch, A, R), device="cuda", generator=g) * 1000 + 10
    flat = torch.rand((16, R), device="cuda", generator=g) * 10 + 1000
    dark = torch.rand((16, R), device="cuda", generator=g) * 5
    out = torch.empty_like(raw)
    normalize(raw, flat, dark, out=out, ctx=ctx)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        normalize(raw, flat, dark, out=out, ctx=ctx)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 8.0 * raw.numel() / (ms * 1e-3) / 1e9
    return {"shape": [batch, A, R], "ms": round(ms, 4), "achieved_GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def _batched_case(ctx, specs, iters: int, seed: int) -> dict:
    """specs: [(N, angles, images)], centred unit bins.  Batched SIRT (one persistent launch
    sequence, one hipGraph) against the same problems group by group through sk_sirt with
    each geometry's kernel shapes autotuned on its batch; each path runs twice and the second (warm) run
    is timed.  Sinograms are uniform random (timing is data-independent)."""
    import torch
    from synkrotomo_amd.projector import centred_bins, sirt_batched, uniform_angles
    gen = torch.Generator(device="cuda").manual_seed(seed)
    groups = []
    updates = 0.0
    for n, a, cnt in specs:
        r, rz = centred_bins(n)
        geo = ctx.geometry(uniform_angles(a), rz, 1.0, r, n)
        geo.autotune(min(cnt, 64))  # tuned on the group's own batch of images
        groups.append((geo, torch.zeros(cnt, n, n, device="cuda"),
                       torch.rand(cnt, a, r, device="cuda", generator=gen)))
        updates += 2.0 * iters * cnt * n * n * a
    singles = [torch.zeros_like(img) for _, img, _ in groups]

    def run_batched():
        for _, img, _ in groups:
            img.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sirt_batched(groups, iters)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def run_groups():
        for img in singles:
            img.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for (geo, _, sino), img in zip(groups, singles):
            geo.sirt_(img, sino, iters)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    run_batched()  # first calls build the plans and capture the graphs
    run_groups()
    t_batched = run_batched()
    t_groups = run_groups()
    same = all(bool(torch.equal(a[1], b)) for a, b in zip(groups, singles))
    del groups, singles
    torch.cuda.empty_cache()
    return {"iterations": iters, "batched_s": round(t_batched, 4), "per_group_s": round(t_groups, 4),
            "batched_gups": round(updates / t_batched / 1e9, 1), "identical_to_per_group": same}


def batched_leg(ctx, iters: int = 30) -> dict:
    """Batched SIRT on two problem sets: large groups (64 x 256^2 / 180 angles, 96 x 512^2 /
    360, 32 x 1024^2 / 720 and a SIRT3D stack of 512 slices of 512^2 / 512), and 16 small
    groups of 4 images each (N = 64, 96, ..., 544 with 0.7 N angles), where one geometry alone
    does not fill the GPU."""
    big = _batched_case(ctx, [(1024, 720, 32), (512, 360, 96), (256, 180, 64), (512, 512, 512)],
                        iters, 30)
    big["problems"] = "64x256^2/180 + 96x512^2/360 + 32x1024^2/720 + SIRT3D 512x512^2/512"
    small = _batched_case(ctx, [(n, int(0.7 * n), 4) for n in range(64, 545, 32)], iters, 31)
    small["problems"] = "16 groups x 4 images, N = 64..544 step 32, 0.7N angles"
    out = dict(big)
    out["small_groups"] = small
    return out


def learned_leg(ctx, n_test: int = 128) -> dict:
    """Time to a data residual, learned SIRT (5 x (SIRT step + BF16 CNN), trained weights in
    synkrotomo_amd/weights) vs plain SIRT, on held-out 512^2 phantoms, 360 angles, 725 bins,
    Poisson noise I0 = 1e4.  Targets: the literal 1e-2 and 1.1x each phantom's noise floor
    (the 1e-2 target lies below the noise floor at I0 = 1e4, so plain SIRT reaches it only by
    fitting noise, after hundreds of iterations or not within the cap of 1500).  The first
    n_test phantoms of the seeded test set; at 128 their median speedup is about the
    1,000-phantom one (5.03-5.14x on two hosts against 5.03x), while the first 64 alone sit
    lower (4.74-5.07x across hosts)."""
    from synkrotomo_amd import learned_run
    if not os.path.exists(learned_run.DEFAULT_WEIGHTS):
        return {"skipped": "no trained weights"}
    out = learned_run.evaluate(ctx, learned_run.DEFAULT_WEIGHTS, n_test, batch=16, max_iters=1500,
                               verbose=False)
    out["conv32"] = conv32_rate(ctx)
    return out


def conv32_rate(ctx, B: int = 16, H: int = 512, W: int = 512, reps: int = 20) -> dict:
    """The 32->32 3x3 BF16 MFMA convolution (k_conv32, bias + ReLU epilogue) on a batch of 16
    512^2 x 32-channel activations: achieved TFLOP/s (2*9*32*32 FLOP per output pixel) against
    the 2.5 PFLOP/s dense BF16 MFMA peak, and the activation bytes it streams against HBM."""
    import torch
    from synkrotomo_amd import _native as nat
    from synkrotomo_amd.projector import _ptr, _stream
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.rand(B, H, W, 32, device="cuda", generator=g).to(torch.bfloat16)
    w = ((torch.rand(32, 3, 3, 32, device="cuda", generator=g) - 0.5) * 0.1).to(torch.bfloat16)
    b = torch.zeros(32, device="cuda")
    y = torch.empty_like(x)
    L, h, s = nat.lib(), ctx.handle, _stream(ctx.device)
    run = lambda: nat.check(L.sk_conv32(h, B, H, W, _ptr(x), None, _ptr(w), _ptr(b), _ptr(y), s), h, "conv32")
    for _ in range(3):
        run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = 2.0 * 9 * 32 * 32 * B * H * W
    act_bytes = 2.0 * 2 * 32 * B * H * W  # bf16 activations in and out
    return {"shape": f"{B}x{H}x{W}x32 bf16", "ms": round(ms, 4),
            "tflops": round(flops / (ms * 1e-3) / 1e12, 1),
            "mfma_peak_tflops": 2500.0, "mfma_frac": round(flops / (ms * 1e-3) / 2.5e15, 4),
            "activation_GBs": round(act_bytes / (ms * 1e-3) / 1e9, 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--exact", action="store_true", help="IEEE division everywhere (bitwise mode)")
    ap.add_argument("--no-sparse", action="store_true", help="skip the explicit-matrix leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-SIRT leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-SIRT leg")
    ap.add_argument("--no-sirt", action="store_true", help="skip the 2048^2 SIRT leg")
    ap.add_argument("--no-stream", action="store_true", help="skip the out-of-core SIRT3D leg")
    ap.add_argument("--dense-only", action="store_true",
                    help="only the dense fp/bp steps (every side leg and the CPU baseline skipped)")
    ap.add_argument("--no-prep", action="store_true", help="skip the normalisation leg")
    ap.add_argument("--only", choices=["sirt", "stream", "prep", "learned", "batched", "sparse", "cpu"],
                    help="the dense steps plus this one side leg (profiling runs)")
    ap.add_argument("--batched-iters", type=int, default=30)
    ap.add_argument("--no-tune", action="store_true", help="use default kernel variants")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device copy first (reference dispatch for HBM counters)")
    args = ap.parse_args()
    if args.dense_only or args.only:
        for leg in ("cpu", "sparse", "batched", "learned", "sirt", "stream", "prep"):
            setattr(args, f"no_{leg}", leg != args.only)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the projector shards by "
              "angle across GPUs, which this build does not run.", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from synkrotomo_amd.phantom import ellipse_phantom
    from synkrotomo_amd.projector import Context, uniform_angles

    ctx = Context.default(0)
    ctx.exact = args.exact
    if args.calibrate_copy:
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    angles = uniform_angles(A)
    phantom = ellipse_phantom(N, PHANTOM_ELLIPSES, seed=PHANTOM_SEED)
    geo = ctx.geometry(angles, RHO0, 1.0, R, N)
    variants = None
    if not args.no_tune:
        from synkrotomo_amd.autotune import tuned
        variants = tuned(geo)
    x = torch.from_numpy(phantom).cuda()
    sino = torch.empty((A, R), device="cuda")
    back = torch.empty((N, N), device="cuda")
    geo.fp(x, out=sino)
    torch.cuda.synchronize()
    sino_host = sino.cpu().numpy()

    for _ in range(args.warmup):
        geo.fp(x, out=sino)
        geo.bp(sino, out=back)
    K = args.steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        ev[k][0].record()
        geo.fp(x, out=sino)
        ev[k][1].record()
        geo.bp(sino, out=back)
        ev[k][2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fp_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / K
    bp_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / K
    upd = float(N) * N * A
    value = 2.0 * upd * K / wall / 1e9
    dom_name, dom_ms = ("bp", bp_ms) if bp_ms >= fp_ms else ("fp", fp_ms)
    roofline = roofline_of(dom_name, {"fp": fp_ms, "bp": bp_ms}, args.exact, variants)

    out = {
        "metric": "GUPS fwd/back @2048²; time-to-residual 1e-2 learned vs plain SIRT",
        "value": round(value, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic {PHANTOM_ELLIPSES}-ellipse phantom, seed {PHANTOM_SEED}",
        "config": {"workload": "dense fp + bp, 2048x2048 image, 2048 angles in [0,pi), 2897 bins, "
                               "deltarho 1, rhozero -1448", "N": N, "angles": A, "bins": R,
                   "mode": "exact" if args.exact else "fast",
                   "variants": {"fp": variants[0], "bp": variants[1]} if variants else "default"},
        "fp_ms": round(fp_ms, 3), "bp_ms": round(bp_ms, 3),
        "fp_gups": round(upd / (fp_ms * 1e-3) / 1e9, 2), "bp_gups": round(upd / (bp_ms * 1e-3) / 1e9, 2),
        "roofline": roofline,
        "cpu_baseline": None,
        "sirt": None,
        "stream3d": None,
        "normalize": None,
        "sparse": None,
        "batched": None,
        "learned": None,
    }
    if not args.no_sirt:
        out["sirt"] = sirt_leg(geo, x, sino)
    if not args.no_stream:
        out["stream3d"] = stream3d_leg(ctx)
    if not args.no_prep:
        out["normalize"] = normalize_leg(ctx)
    if not args.no_learned:
        out["learned"] = learned_leg(ctx)
    if not args.no_batched:
        out["batched"] = batched_leg(ctx, args.batched_iters)
    if not args.no_sparse:
        out["sparse"] = sparse_leg(geo, x, sino, max(3, K // 2))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(angles, phantom.ravel(), sino_host.ravel())
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
