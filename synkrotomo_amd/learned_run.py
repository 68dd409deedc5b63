"""Train the learned-SIRT correction and measure time to a data residual.

    python -m synkrotomo_amd.learned_run train [--phantoms 20000] [--batch 8] [--out PATH]
    python -m synkrotomo_amd.learned_run eval  [--phantoms 1000] [--weights PATH]

Configuration: 512x512 ellipse phantoms (10 ellipses, seeded), 360 angles in [0, pi),
725 bins, Poisson transmission noise with I0 = 1e4 on exp(-s*Ax), 5 SIRT steps each
followed by the correction.  Training phantoms come from one seeded stream and the
held-out test phantoms from another.

Attenuation scale s: fixed so that the 95th percentile of the per-phantom maximum line
integral maps to s*Ax = 3 (exp(-3) = 5% transmission).  With Poisson counts the relative
noise of a ray with s*Ax = t is sqrt(exp(t)/I0)/t, whose minimum over t is e/(2 sqrt(I0))
= 1.36% at I0 = 1e4: no choice of s puts the noise floor ||b - A x_true||/||b|| below 1e-2,
so a data residual of 1e-2 is reached only by fitting noise.  The evaluation therefore
reports two targets per test phantom: the fixed 1e-2, and 1.1x that phantom's own noise
floor (as consistent with the data as the true image, within 10%).

Time to a target = (iterations or stages needed) x (measured time per iteration or stage
on the GPU, batch of images); errors are the relative image error against the true
phantom computed in FP64.  The learned time is the measured time of the whole learned
reconstruction (stages, the finiteness guard, and the equal-cost fallback rule's plain SIRT
steps and residuals) plus the relaxed SIRT iterations after it; the stages-only time is
reported beside it, and so is the residual-judged method (``LearnedSIRT.solve``: the stages
with the finiteness guard, then the iterations to the target, plain SIRT only for an image
that misses it within the budget).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

from .learned import Correction, LearnedSIRT, gpu_phantoms, noisy_sinogram, residual
from .projector import Context, centred_bins, uniform_angles

N, A, STAGES, I0, ELLIPSES = 512, 360, 5, 1e4, 10
TRAIN_SEED, TEST_SEED, NOISE_SEED, VALID_SEED = 2025, 777, 99, 555
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "learned_sirt_512_360.npz")


def geometry(ctx, batch: int = 16):
    """The 512^2 / 360-angle / 725-bin geometry with its kernel shapes autotuned on batches of
    `batch` images (plain and learned SIRT run the same projector kernels on the same batches)."""
    from .autotune import tuned
    R, rz = centred_bins(N)
    geo = ctx.geometry(uniform_angles(A), rz, 1.0, R, N)
    tuned(geo, batch=batch)
    return geo


def attenuation_scale(ctx, geo, rng) -> float:
    x = gpu_phantoms(ctx, rng, 64, N, ELLIPSES)
    b = geo.fp(x)
    peak = b.reshape(64, -1).amax(dim=1).cpu().numpy()
    return float(3.0 / np.percentile(peak, 95))


def train(args) -> int:
    ctx = Context.default(0)
    geo = geometry(ctx, args.batch)
    rng = np.random.default_rng(TRAIN_SEED)
    scale = attenuation_scale(ctx, geo, np.random.default_rng(TRAIN_SEED + 1))
    gen = torch.Generator(device="cuda").manual_seed(NOISE_SEED)
    net = Correction(ctx, seed=0, cin=args.cin)
    if args.init:
        net.load(args.init)  # fine-tune from earlier weights
    model = LearnedSIRT(geo, net, args.stages)
    steps = args.epochs * (args.phantoms // args.batch)
    opt = torch.optim.Adam(net.parameters(), lr=args.lr)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=args.lr, total_steps=steps, pct_start=0.05)
    t0 = time.perf_counter()
    log = []
    skipped = 0
    for step in range(steps):
        if step % (args.phantoms // args.batch) == 0:
            rng = np.random.default_rng(TRAIN_SEED)  # each epoch revisits the same phantoms
        x_true = gpu_phantoms(ctx, rng, args.batch, N, ELLIPSES)
        clean = geo.fp(x_true)
        b = noisy_sinogram(clean, scale, I0, gen)
        stages = model.unrolled(b, checkpoint=True, all_stages=True)
        x = stages[-1]
        loss = torch.mean((x - x_true) ** 2)
        if args.proj_weight > 0:
            # projection-domain error against the noise-free data: ||A(x - x_true)||^2 /
            # ||A x_true||^2 per image (the part of the data residual that is not noise)
            from .autograd import project
            cc = clean.reshape(b.shape[0], -1)
            d = project(x, geo).reshape(b.shape[0], -1) - cc
            # a phantom whose ellipses all clip to zero has no projection: no term for it
            den = cc.square().sum(1)
            loss = loss + args.proj_weight * torch.mean(torch.where(den > 0, d.square().sum(1) / den.clamp_min(1e-30),
                                                                    torch.zeros_like(den)))
        if args.stage_weight > 0:
            for xs in stages[:-1]:
                loss = loss + args.stage_weight * torch.mean((xs - x_true) ** 2)
        if args.tail_steps > 0 and args.tail_weight > 0:
            # the data residual after `tail_steps` relaxed SIRT steps from the last stage, above
            # 1.1x the phantom's noise floor: the tail is affine in the stage output, so this
            # rewards outputs whose remaining error the tail removes fastest
            from .autograd import project
            xt = x
            for _ in range(args.tail_steps):
                xt = xt + args.tail_relax * (model.sirt_step(xt, b) - xt)
            bb = b.reshape(b.shape[0], -1)
            bn = bb.square().sum(1)
            r2 = (project(xt, geo).reshape(b.shape[0], -1) - bb).square().sum(1) / bn
            with torch.no_grad():
                f2 = (clean.reshape(b.shape[0], -1) - bb).square().sum(1) / bn
            ok = f2 > 0
            excess = torch.where(ok, torch.relu(r2 - 1.21 * f2) / f2.clamp_min(1e-30), torch.zeros_like(r2))
            loss = loss + args.tail_weight * torch.mean(excess)
        if args.res_weight > 0 or args.res_excess_weight > 0:
            # data consistency of the final image: ||A x - b|| / ||b|| per image
            from .autograd import project
            bb = b.reshape(b.shape[0], -1)
            bn = bb.square().sum(1)
            r2 = (project(x, geo).reshape(b.shape[0], -1) - bb).square().sum(1) / bn
            if args.res_weight > 0:
                loss = loss + args.res_weight * torch.mean(r2)
            if args.res_excess_weight > 0:
                # only the part of the residual above 1.1x the phantom's own noise floor
                with torch.no_grad():
                    floor = ((geo.fp(x_true).reshape(b.shape[0], -1) - bb).square().sum(1) / bn).sqrt()
                excess = torch.relu(r2.clamp_min(1e-20).sqrt() - 1.1 * floor)
                loss = loss + args.res_excess_weight * torch.mean(excess.square())
        opt.zero_grad(set_to_none=True)
        if not torch.isfinite(loss):
            # a non-finite loss would poison Adam's moments: skip the batch, keep the schedule
            skipped += 1
            sched.step()
            continue
        loss.backward()
        grads_ok = torch.stack([torch.isfinite(q.grad).all() for q in net.parameters() if q.grad is not None]).all()
        if not bool(grads_ok):
            skipped += 1
            sched.step()
            continue
        opt.step()
        sched.step()
        if step % 100 == 0 or step == steps - 1:
            torch.cuda.synchronize()
            rec = {"step": step, "loss": float(loss.detach()), "skipped": skipped,
                   "elapsed_s": round(time.perf_counter() - t0, 1)}
            log.append(rec)
            print(json.dumps(rec), flush=True)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    net.save(args.out, scale=scale, I0=I0, stages=args.stages, N=N, A=A, phantoms=args.phantoms,
             batch=args.batch, epochs=args.epochs, res_weight=args.res_weight,
             stage_weight=args.stage_weight, res_excess_weight=args.res_excess_weight,
             proj_weight=args.proj_weight, tail_steps=args.tail_steps, tail_weight=args.tail_weight,
             init=os.path.basename(args.init) if args.init else "",
             lr=args.lr, train_seconds=train_s)
    print(json.dumps({"saved": args.out, "train_seconds": round(train_s, 1), "scale": scale}), flush=True)
    return 0


def _first_hit(ks: np.ndarray, res: np.ndarray, target: np.ndarray) -> np.ndarray:
    """res [records, B] recorded after ks[i] iterations -> per image the first recorded
    iteration count with res <= target, or -1."""
    hit = res <= target[None, :]
    return np.where(hit.any(axis=0), ks[hit.argmax(axis=0)], -1)


def _track(step, x, geo, b, x_true, start: int, max_iters: int, dense: int, stride: int, stop: float):
    """Iterate step(x) (one SIRT step in place) from iteration `start` to `max_iters`, recording
    the data residual and the FP64 image error of every image after iteration k for k <= dense
    and every `stride` iterations beyond (and at the end).  Stops once every image's residual
    is at or below `stop`.  Returns (ks, res [records, B], err [records, B])."""
    B = x.shape[0]
    xt_norm = x_true.reshape(B, -1).double().norm(dim=1)
    ks, res, err = [], [], []

    def record(k):
        ks.append(k)
        res.append(residual(geo, x, b).cpu().numpy())
        err.append(((x - x_true).reshape(B, -1).double().norm(dim=1) / xt_norm).cpu().numpy())

    record(start)
    k = start
    while k < max_iters and float(res[-1].max()) > stop:
        step(x)
        k += 1
        if k <= dense or k % stride == 0 or k == max_iters:
            record(k)
    return np.asarray(ks), np.asarray(res), np.asarray(err)


def evaluate(ctx, weights: str, n_test: int, batch: int = 16, max_iters: int = 300,
             verbose: bool = True, dense: int = 300, stride: int = 10,
             tail_relax: float | None = None, seed: int = TEST_SEED) -> dict:
    """Time to a data residual, learned vs plain SIRT, on `n_test` held-out phantoms.  Plain SIRT
    (the reference's, relaxation 1) runs from zero; the learned method runs its stages, then SIRT
    steps relaxed by its tail factor (the weights' `tail_relax`, or the argument).  Residuals and
    image errors are recorded after every iteration up to `dense` and every `stride` iterations
    beyond, up to `max_iters` or until every image is at or below 1e-2: an iteration count past
    `dense` is an upper bound within `stride`."""
    geo = geometry(ctx, batch)
    net = Correction(ctx)
    meta = net.load(weights)
    scale = float(meta["scale"])
    stages = int(meta["stages"]) if "stages" in meta else STAGES
    if tail_relax is None:
        tail_relax = float(meta["tail_relax"]) if "tail_relax" in meta else 1.0
    model = LearnedSIRT(geo, net, stages, tail_relax=tail_relax)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda").manual_seed(NOISE_SEED + 1)
    # timing of one SIRT iteration, one relaxed tail iteration and one stage (SIRT step +
    # correction), batch of images
    xt = gpu_phantoms(ctx, rng, batch, N, ELLIPSES)
    b = noisy_sinogram(geo.fp(xt), scale, I0, gen)
    xb = torch.zeros_like(xt)
    with torch.no_grad():
        for _ in range(3):
            geo.sirt_(xb, b, 1)
            model.tail_(xb, b, 1)
            model.stage(xb, b, 1)
        model.reconstruct(b)  # warm-up
        model.reconstruct(b, guard=False)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        e[0].record()
        geo.sirt_(xb, b, 20)
        e[1].record()
        for _ in range(5):
            model.stage(xb, b, 1)
        e[2].record()
        # the whole learned reconstruction as used: stages with the finiteness guard, the
        # plain SIRT steps of the fallback rule and both residuals
        model.reconstruct(b)
        e[3].record()
        model.tail_(xb, b, 20)
        e[4].record()
        # the stages with the finiteness guard only (the residual-judged method: the data
        # residual decides, plain SIRT is the fallback if the target is not reached)
        model.reconstruct(b, guard=False)
        e[5].record()
        torch.cuda.synchronize()
    t_iter = e[0].elapsed_time(e[1]) / 20 / batch * 1e-3
    t_stage = e[1].elapsed_time(e[2]) / 5 / batch * 1e-3
    t_recon = e[2].elapsed_time(e[3]) / batch * 1e-3
    t_tail = e[3].elapsed_time(e[4]) / 20 / batch * 1e-3
    t_recon_ng = e[4].elapsed_time(e[5]) / batch * 1e-3

    rows = []
    for start in range(0, n_test, batch):
        B = min(batch, n_test - start)
        x_true = gpu_phantoms(ctx, rng, B, N, ELLIPSES)
        b = noisy_sinogram(geo.fp(x_true), scale, I0, gen)
        floor = residual(geo, x_true, b).cpu().numpy()
        with torch.no_grad():
            x = torch.zeros_like(x_true)
            kp, res_p, err_p = _track(lambda z: geo.sirt_(z, b, 1), x, geo, b, x_true, 0, max_iters,
                                      dense, stride, 1e-2)
            xl, fell = model.reconstruct(b)
            kl, res_l, err_l = _track(lambda z: model.tail_(z, b, 1), xl, geo, b, x_true, 0,
                                      max_iters - stages, dense, stride, 1e-2)
            if bool(fell.any()):  # the unguarded stages differ from the guarded output there
                xn, _ = model.reconstruct(b, guard=False)
                kn, res_n, err_n = _track(lambda z: model.tail_(z, b, 1), xn, geo, b, x_true, 0,
                                          max_iters - stages, dense, stride, 1e-2)
            else:
                kn, res_n, err_n = kl, res_l, err_l
        k5 = int(np.searchsorted(kp, stages))
        for tname, target in (("1e-2", np.full(B, 1e-2)), ("1.1x_floor", 1.1 * floor)):
            hp = _first_hit(kp, res_p, target)
            hl = _first_hit(kl, res_l, target)
            hn = _first_hit(kn, res_n, target)
            jn = np.array([int(np.searchsorted(kn, h)) if h >= 0 else -1 for h in hn])
            # image error plain SIRT has when it reaches the target, and the first learned
            # state (after the stages, then after tail iterations) at or below it
            ip = np.array([int(np.searchsorted(kp, h)) if h >= 0 else -1 for h in hp])
            il = np.array([int(np.searchsorted(kl, h)) if h >= 0 else -1 for h in hl])
            e_at = np.array([err_p[ip[i], i] if ip[i] >= 0 else -1.0 for i in range(B)])
            he = _first_hit(kl, err_l, e_at)
            for i in range(B):
                rows.append({"target": tname, "floor": float(floor[i]), "fell_back": bool(fell[i]),
                             "plain_iters": int(hp[i]),
                             "plain_err": float(err_p[ip[i], i]) if hp[i] >= 0 else None,
                             "learned_extra_iters": int(hl[i]),
                             "learned_err": float(err_l[il[i], i]) if hl[i] >= 0 else None,
                             "plain_res_at_end": float(res_p[-1, i]), "plain_iters_run": int(kp[-1]),
                             "learned_res_after_stages": float(res_l[0, i]),
                             "learned_err_after_stages": float(err_l[0, i]),
                             "plain_err_at_stages": float(err_p[min(k5, len(kp) - 1), i]),
                             "learned_extra_iters_to_plain_err": int(he[i]) if hp[i] >= 0 else -1,
                             "judged_extra_iters": int(hn[i]),
                             "judged_err": float(err_n[jn[i], i]) if hn[i] >= 0 else None})
        if verbose:
            print(json.dumps({"done": start + B}), flush=True)

    out = {"t_iter_ms": round(t_iter * 1e3, 4), "t_tail_iter_ms": round(t_tail * 1e3, 4),
           "t_stage_ms": round(t_stage * 1e3, 4), "t_reconstruct_ms": round(t_recon * 1e3, 4),
           "t_stages_unguarded_ms": round(t_recon_ng * 1e3, 4),
           "scale": scale, "tail_relax": tail_relax,
           "stage_relax": [round(float(v), 4) for v in net.relax.detach().cpu()[1:stages]],
           "n_test": n_test, "max_iters": max_iters, "dense_iters": dense, "stride": stride,
           "noise_floor_median": float(np.median([r["floor"] for r in rows if r["target"] == "1e-2"]))}
    for tname in ("1e-2", "1.1x_floor"):
        rs = [r for r in rows if r["target"] == tname]
        both = [r for r in rs if r["plain_iters"] >= 0 and r["learned_extra_iters"] >= 0]
        summary = {"reached_plain": sum(r["plain_iters"] >= 0 for r in rs),
                   "reached_learned": sum(r["learned_extra_iters"] >= 0 for r in rs), "of": len(rs)}
        if tname == "1e-2":
            summary["median_plain_res_at_end"] = round(float(np.median([r["plain_res_at_end"] for r in rs])), 5)
            summary["plain_iters_run_max"] = max(r["plain_iters_run"] for r in rs)
        if both:
            tp = np.array([r["plain_iters"] * t_iter for r in both])
            tl = np.array([t_recon + r["learned_extra_iters"] * t_tail for r in both])
            ts = np.array([stages * t_stage + r["learned_extra_iters"] * t_tail for r in both])
            summary.update({
                "median_plain_ms": round(float(np.median(tp)) * 1e3, 3),
                "median_learned_ms": round(float(np.median(tl)) * 1e3, 3),
                "speedup_median": round(float(np.median(tp) / np.median(tl)), 2),
                "median_learned_stages_only_ms": round(float(np.median(ts)) * 1e3, 3),
                "speedup_stages_only_median": round(float(np.median(tp) / np.median(ts)), 2),
                "median_plain_iters": float(np.median([r["plain_iters"] for r in both])),
                "median_learned_extra_iters": float(np.median([r["learned_extra_iters"] for r in both])),
                "median_plain_err": round(float(np.median([r["plain_err"] for r in both])), 5),
                "median_learned_err": round(float(np.median([r["learned_err"] for r in both])), 5),
                "learned_err_le_plain": sum(r["learned_err"] <= r["plain_err"] for r in both)})
            # the same image error as plain SIRT at its target, reached by the learned method
            eq = [r for r in rs if r["plain_iters"] >= 0 and r["learned_extra_iters_to_plain_err"] >= 0]
            if eq:
                tpe = np.array([r["plain_iters"] * t_iter for r in eq])
                tle = np.array([t_recon + r["learned_extra_iters_to_plain_err"] * t_tail for r in eq])
                summary["equal_image_error"] = {
                    "of": len(eq), "median_learned_extra_iters": float(np.median(
                        [r["learned_extra_iters_to_plain_err"] for r in eq])),
                    "median_learned_ms": round(float(np.median(tle)) * 1e3, 3),
                    "speedup_median": round(float(np.median(tpe) / np.median(tle)), 2)}
            # residual-judged method (LearnedSIRT.solve): the stages with the finiteness guard,
            # then tail steps to the target; plain SIRT only for an image that misses it
            jb = [r for r in rs if r["plain_iters"] >= 0 and r["judged_extra_iters"] >= 0]
            if jb:
                tpj = np.array([r["plain_iters"] * t_iter for r in jb])
                tlj = np.array([t_recon_ng + r["judged_extra_iters"] * t_tail for r in jb])
                summary["residual_judged"] = {
                    "of": len(jb), "missed_target": sum(r["judged_extra_iters"] < 0 for r in rs),
                    "median_learned_ms": round(float(np.median(tlj)) * 1e3, 3),
                    "speedup_median": round(float(np.median(tpj) / np.median(tlj)), 2),
                    "median_learned_extra_iters": float(np.median([r["judged_extra_iters"] for r in jb])),
                    "median_learned_err": round(float(np.median([r["judged_err"] for r in jb])), 5),
                    "median_plain_err": round(float(np.median([r["plain_err"] for r in jb])), 5)}
        out[tname] = summary
    out["fell_back"] = sum(r["fell_back"] for r in rows if r["target"] == "1e-2")
    rs = [r for r in rows if r["target"] == "1e-2"]
    out["stages"] = stages
    out["after_stages"] = {"median_learned_err": round(float(np.median([r["learned_err_after_stages"] for r in rs])), 5),
                      "median_plain_err": round(float(np.median([r["plain_err_at_stages"] for r in rs])), 5),
                      "median_learned_res": round(float(np.median([r["learned_res_after_stages"] for r in rs])), 5)}
    return out


def tune_tail(ctx, weights: str, n: int, candidates, batch: int = 16, max_iters: int = 300) -> dict:
    """Pick the tail relaxation on validation phantoms (a seed apart from the training and test
    streams): the factor with the shortest median learned time to 1.1x the noise floor whose
    median image error there is not above plain SIRT's.  Stored in the weights' metadata."""
    results = {}
    for lam in candidates:
        out = evaluate(ctx, weights, n, batch=batch, max_iters=max_iters, verbose=False,
                       tail_relax=lam, seed=VALID_SEED)
        t = out["1.1x_floor"]
        results[lam] = t
        print(json.dumps({"tail_relax": lam, "speedup_median": t.get("speedup_median"),
                          "median_learned_err": t.get("median_learned_err"),
                          "median_plain_err": t.get("median_plain_err"),
                          "median_learned_extra_iters": t.get("median_learned_extra_iters")}), flush=True)
    ok = [l for l, t in results.items() if t.get("reached_learned") == t.get("of") and
          t.get("median_learned_err", 1e9) <= t.get("median_plain_err", -1)]
    best = max(ok, key=lambda l: results[l]["speedup_median"]) if ok else 1.0
    z = np.load(weights, allow_pickle=False)
    arrs = {k: z[k] for k in z.files}
    arrs["meta_tail_relax"] = np.asarray(best)
    np.savez(weights, **arrs)
    return {"tail_relax": best, "candidates": {str(k): v.get("speedup_median") for k, v in results.items()}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--phantoms", type=int, default=20000)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--epochs", type=int, default=3)
    t.add_argument("--stages", type=int, default=STAGES, help="SIRT step + correction stages")
    t.add_argument("--res-weight", type=float, default=0.0,
                   help="weight of the final data residual ||Ax-b||^2/||b||^2 in the loss")
    t.add_argument("--stage-weight", type=float, default=0.0,
                   help="weight of the image error after each earlier stage")
    t.add_argument("--res-excess-weight", type=float, default=0.0,
                   help="weight of the squared excess of the final residual over 1.1x the "
                        "phantom's noise floor")
    t.add_argument("--proj-weight", type=float, default=0.0,
                   help="weight of the final image's projection error against the noise-free "
                        "data, ||A(x - x_true)||^2 / ||A x_true||^2")
    t.add_argument("--init", default="", help="start from these weights instead of a fresh init")
    t.add_argument("--tail-steps", type=int, default=0,
                   help="relaxed SIRT steps after the stages whose final residual enters the loss")
    t.add_argument("--tail-weight", type=float, default=0.0,
                   help="weight of that residual's excess over 1.1x the noise floor (relative)")
    t.add_argument("--tail-relax", type=float, default=1.8, help="relaxation of those steps")
    t.add_argument("--cin", type=int, default=1, choices=[1, 2],
                   help="network input channels: 1 = the image after each SIRT step, 2 = also the "
                        "step's update direction")
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    e = sub.add_parser("eval")
    e.add_argument("--phantoms", type=int, default=1000)
    e.add_argument("--batch", type=int, default=16)
    e.add_argument("--max-iters", type=int, default=3000)
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--tail-relax", type=float, default=None,
                   help="relaxation of the SIRT steps after the stages (default: the weights')")
    tt = sub.add_parser("tune-tail")
    tt.add_argument("--phantoms", type=int, default=64)
    tt.add_argument("--weights", default=DEFAULT_WEIGHTS)
    tt.add_argument("--candidates", default="1.0,1.5,1.7,1.8,1.9")
    args = ap.parse_args(argv)
    if args.cmd == "train":
        return train(args)
    if args.cmd == "tune-tail":
        res = tune_tail(Context.default(0), args.weights, args.phantoms,
                        [float(v) for v in args.candidates.split(",")])
    else:
        res = evaluate(Context.default(0), args.weights, args.phantoms, args.batch, args.max_iters,
                       tail_relax=args.tail_relax)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
