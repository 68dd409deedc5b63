"""Build libsynkro.so (the HIP kernels + C ABI) in-tree for gfx950.

    python -m synkrotomo_amd.build

Each csrc/*.hip file is compiled to an object in parallel, then linked into
synkrotomo_amd/lib/libsynkro.so; probes/*.hip (stand-alone measurement programs, not part of
the library) become executables of the same name in synkrotomo_amd/lib/.  Flags:
  -ffp-contract=off       every multiply and add stays a separate IEEE operation (the
                          projector's pixel indices depend on it)
  -packed-fp32-ops off    no packed FP32 instructions are generated
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libsynkro.so")
PROBES = os.path.join(PKG, "probes")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SYNKRO_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def _probe(src: str) -> str:
    exe = os.path.join(LIBDIR, os.path.splitext(os.path.basename(src))[0])
    cmd = [HIPCC, *CFLAGS, src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return exe


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    probes = sorted(glob.glob(os.path.join(PROBES, "*.hip")))
    # at most 16 compiles at once: a host's core count can exceed its share of it
    workers = min(len(srcs) + len(probes), 16, max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        probe_jobs = [ex.submit(_probe, p) for p in probes]
        objs = list(ex.map(_compile, srcs))
        for j in probe_jobs:
            j.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
