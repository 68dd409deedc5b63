"""SIRT3D over slice stacks held in host memory, streamed through the GPU in chunks.

The reference's SIRT3D maps SIRT over the slices of one geometry with the whole stack on the
device (futhark/SIRT3D.fut:18-28).  ``sirt3d_streamed`` runs the same per-slice SIRT on stacks
of any size: chunks of slices are uploaded, reconstructed and downloaded on three HIP streams
(default priority), double-buffered on both sides, so the upload of chunk k+1 and the download
of chunk k-1 run while chunk k is reconstructed.  Host staging is pinned memory; the caller's
arrays may be pageable.  Each slice gets exactly the kernels it gets in core (SIRT from a zero
image; batches are independent slices), so results are bitwise those of the in-core run.

    out = sirt3d_streamed(geo, projections, iterations, chunk_slices=64)
    # projections: host float32 [S, A, R] (or flat); out: host float32 [S, N, N]
"""
from __future__ import annotations

import numpy as np
import torch

from .projector import Geometry


def sirt3d_streamed(geo: Geometry, projections, iterations: int, chunk_slices: int,
                    out: np.ndarray | None = None) -> np.ndarray:
    if chunk_slices < 1:
        raise ValueError("chunk_slices must be >= 1")
    if int(iterations) < 0:
        raise ValueError("iterations must be >= 0")
    A, R, N = geo.A, geo.R, geo.N
    proj = np.asarray(projections)
    if proj.dtype != np.float32:
        raise TypeError("projections must be float32")
    if proj.size % (A * R):
        raise ValueError(f"projections hold {proj.size} values, not a multiple of A*R = {A * R}")
    S = proj.size // (A * R)
    proj = proj.reshape(S, A, R)
    if out is None:
        out = np.empty((S, N, N), np.float32)
    elif out.dtype != np.float32 or out.size != S * N * N or not out.flags.c_contiguous:
        raise ValueError("out must be a C-contiguous float32 array of S*N*N values")
    out3 = out.reshape(S, N, N)
    if S == 0:
        return out
    C = min(chunk_slices, S)
    dev = torch.device("cuda", geo.ctx.device)
    pin_in = [torch.empty((C, A, R), dtype=torch.float32, pin_memory=True) for _ in range(2)]
    pin_out = [torch.empty((C, N, N), dtype=torch.float32, pin_memory=True) for _ in range(2)]
    d_sino = [torch.empty((C, A, R), dtype=torch.float32, device=dev) for _ in range(2)]
    d_img = [torch.empty((C, N, N), dtype=torch.float32, device=dev) for _ in range(2)]
    s_in, s_run, s_out = (torch.cuda.Stream(device=dev) for _ in range(3))
    ev_in = [torch.cuda.Event() for _ in range(2)]      # upload of a buffer pair finished
    ev_run = [torch.cuda.Event() for _ in range(2)]     # reconstruction finished
    ev_out = [torch.cuda.Event() for _ in range(2)]     # download finished
    used_in = [False, False]
    used_out = [False, False]
    chunks = [(k, lo, min(lo + C, S)) for k, lo in enumerate(range(0, S, C))]

    def finish(k: int, lo: int, hi: int) -> None:
        b = k & 1
        ev_out[b].synchronize()
        out3[lo:hi] = pin_out[b][: hi - lo].numpy()

    torch.cuda.current_stream(dev).synchronize()  # buffers allocated on the current stream
    for k, lo, hi in chunks:
        b, n = k & 1, hi - lo
        if used_in[b]:
            ev_in[b].synchronize()          # pinned buffer b's previous upload has finished
        pin_in[b][:n].numpy()[...] = proj[lo:hi]
        with torch.cuda.stream(s_in):
            if used_in[b]:
                s_in.wait_event(ev_run[b])  # device buffer b's previous chunk is reconstructed
            d_sino[b][:n].copy_(pin_in[b][:n], non_blocking=True)
            ev_in[b].record(s_in)
        used_in[b] = True
        with torch.cuda.stream(s_run):
            s_run.wait_event(ev_in[b])
            if used_out[b]:
                s_run.wait_event(ev_out[b])  # image buffer b's previous chunk is downloaded
            img = d_img[b][:n]
            img.zero_()
            geo.sirt_(img, d_sino[b][:n], int(iterations))
            ev_run[b].record(s_run)
        if used_out[b]:
            finish(k - 2, *chunks[k - 2][1:])   # chunk k-2 used pinned output buffer b
        with torch.cuda.stream(s_out):
            s_out.wait_event(ev_run[b])
            pin_out[b][:n].copy_(d_img[b][:n], non_blocking=True)
            ev_out[b].record(s_out)
        used_out[b] = True
    for k, lo, hi in chunks[-2:]:
        finish(k, lo, hi)
    for s in (s_in, s_run, s_out):
        s.synchronize()
    return out
