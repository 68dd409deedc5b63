"""Futhark's value formats: read and write the reference's datasets.

The reference benchmarks its entry points with ``futhark bench`` on dataset files named in
each program's header (``-- input@data/fpinputf32rad512``, futhark/forwardprojection.fut:1-8,
futhark/SIRT3D.fut:1-14).  A dataset is the entry point's arguments in order, as a stream of
values, each in one of Futhark's two published value formats:

* text: scalars with a type suffix (``725i32``, ``-362.0f32``; no suffix: ``i32`` for
  integers, ``f64`` for decimals), ``true``/``false``, ``f32.nan``, ``f32.inf``,
  ``-f32.inf``; arrays ``[v, v, ...]``, nested for more dimensions; empty arrays
  ``empty([0][3]f32)``; ``--`` starts a comment.
* binary: the byte ``b``, the format version (2), the rank (one byte), the element type as
  four ASCII bytes (``" f32"``, ``"  u8"``, ``"bool"``, ...), the shape as rank little-endian
  int64s, then the elements, little-endian, row-major (bool: one byte each).

``load``/``loads`` read any mix of both; ``dump``/``dumps`` write either.  Scalars come back
as numpy scalars, arrays as numpy arrays of the element type.  Text decimals are read as the
nearest double, then rounded to the element type: the reference's files print doubles in
full (``0.7853981633974483f32``), so that is one rounding.  The text writer prints each
float as the shortest decimal of its exact double value, so text round trips are exact.

Command line (the ``futhark run`` / ``futhark bench`` of this package)::

    python -m synkrotomo_amd.futhark_data run ENTRY DATASET [-o OUT] [--binary]
    python -m synkrotomo_amd.futhark_data bench ENTRY DATASET [--runs 10]

ENTRY is a drop-in module name (forwardprojection, backprojection, their _sparse
variants, SIRT, SIRT3D).  ``run`` writes the result as a value (to OUT or stdout); ``bench``
prints the mean wall time per call in microseconds, output transfer included, as the
reference's ``size;us`` result files do (output/*/fut_fp.csv).
"""
from __future__ import annotations

import argparse
import json
import re
import struct
import sys
import time
from typing import Iterable

import numpy as np

# Futhark element type <-> numpy dtype (little-endian), and the binary type field.
_TYPES = {
    "i8": np.dtype("<i1"), "i16": np.dtype("<i2"), "i32": np.dtype("<i4"), "i64": np.dtype("<i8"),
    "u8": np.dtype("<u1"), "u16": np.dtype("<u2"), "u32": np.dtype("<u4"), "u64": np.dtype("<u8"),
    "f16": np.dtype("<f2"), "f32": np.dtype("<f4"), "f64": np.dtype("<f8"), "bool": np.dtype("?"),
}
_BIN_NAME = {t: t.rjust(4).encode("ascii") for t in _TYPES}
_BIN_TYPE = {v: k for k, v in _BIN_NAME.items()}
BINARY_VERSION = 2

_SUFFIX = re.compile(r"(i8|i16|i32|i64|u8|u16|u32|u64|f16|f32|f64)$")
_TOKEN = re.compile(r"--[^\n]*|\[|\]|,|empty\s*\(|\)|[^\s\[\],()]+")


def type_name(dtype) -> str:
    """Futhark name of a numpy dtype (``float32`` -> ``f32``)."""
    d = np.dtype(dtype).newbyteorder("<") if np.dtype(dtype).kind != "b" else np.dtype("?")
    for k, v in _TYPES.items():
        if v == d:
            return k
    raise TypeError(f"no Futhark type for {dtype}")


# ---------------------------------------------------------------------------- text reader
def _scalar(tok: str):
    """(type name, python value) of one text literal."""
    if tok in ("true", "false"):
        return "bool", tok == "true"
    neg = tok.startswith("-")
    body = tok[1:] if neg or tok.startswith("+") else tok
    m = re.fullmatch(r"(f16|f32|f64)\.(nan|inf)", body)
    if m:
        v = float("nan") if m.group(2) == "nan" else float("inf")
        return m.group(1), -v if neg else v
    m = _SUFFIX.search(body)
    t = m.group(1) if m else None
    num = (body[: m.start()] if m else body).replace("_", "")
    if not num:
        raise ValueError(f"bad literal {tok!r}")
    is_hex = num[:2].lower() in ("0x",)
    is_float = (not is_hex and any(c in num for c in ".eE")) or (is_hex and "p" in num.lower())
    if t is None:
        t = "f64" if is_float else "i32"
    if t[0] == "f":
        v = float.fromhex(num) if is_hex else float(num)
    else:
        if is_float:
            raise ValueError(f"integer type with a fractional literal {tok!r}")
        v = int(num, 0)
    return t, -v if neg else v


def _make(t: str, v):
    return _TYPES[t].type(v)


class _Parser:
    def __init__(self, text: str):
        self.toks = [t for t in _TOKEN.findall(text) if not t.startswith("--")]
        self.i = 0

    def done(self) -> bool:
        return self.i >= len(self.toks)

    def _next(self) -> str:
        if self.i >= len(self.toks):
            raise ValueError("unexpected end of input")
        t = self.toks[self.i]
        self.i += 1
        return t

    def value(self):
        t = self._next()
        if t == "[":
            return self._array()
        if t.startswith("empty"):
            return self._empty()
        if t in ("]", ",", ")"):
            raise ValueError(f"unexpected {t!r}")
        ty, v = _scalar(t)
        return _make(ty, v)

    def _flat_fast(self):
        """A one-dimensional array of plain decimal literals with one suffix, converted in
        bulk (the reference's dataset files are such arrays); None if it is not one."""
        toks, i = self.toks, self.i
        try:
            k = toks.index("]", i)
        except ValueError:
            return None
        body = toks[i:k]
        if not body or len(body) % 2 != 1 or "[" in body or any(t.startswith("empty") for t in body):
            return None
        elems, seps = body[0::2], body[1::2]
        if any(c != "," for c in seps):
            return None
        suf = elems[0][-3:]
        if suf not in ("f32", "f64", "i32", "i64") or any(e[-3:] != suf for e in elems):
            return None
        nums = [e[:-3] for e in elems]
        try:
            if suf[0] == "f":
                a = np.array([float(n) for n in nums], dtype=np.float64).astype(_TYPES[suf])
            else:
                a = np.array([int(n, 10) for n in nums], dtype=_TYPES[suf])
        except ValueError:
            return None
        self.i = k + 1
        return a

    def _array(self):
        elems, types, nested = [], set(), False
        if self.i >= len(self.toks) or self.toks[self.i] == "]":
            raise ValueError("an empty array is written empty([0]t)")
        fast = self._flat_fast()
        if fast is not None:
            return fast
        while True:
            t = self._next()
            if t == "[":
                elems.append(self._array())
                nested = True
            elif t.startswith("empty"):
                elems.append(self._empty())
                nested = True
            else:
                ty, v = _scalar(t)
                types.add(ty)
                elems.append(v)
            sep = self._next()
            if sep == "]":
                break
            if sep != ",":
                raise ValueError(f"expected ',' or ']', got {sep!r}")
        if nested:
            if not all(isinstance(e, np.ndarray) for e in elems):
                raise ValueError("array mixes scalars and arrays")
            shapes = {e.shape for e in elems}
            dts = {e.dtype for e in elems}
            if len(shapes) != 1 or len(dts) != 1:
                raise ValueError("irregular array")
            return np.stack(elems)
        if len(types) != 1:
            raise ValueError(f"array elements of different types {sorted(types)}")
        return np.array(elems, dtype=_TYPES[types.pop()])

    def _empty(self):
        # empty([d0][d1]...t) -- the tokenizer split it at the brackets
        dims, t = [], None
        while True:
            tok = self._next()
            if tok == ")":
                break
            if tok == "[":
                dims.append(int(self._next()))
                if self._next() != "]":
                    raise ValueError("bad empty() shape")
            else:
                t = tok
        if t not in _TYPES or not dims or 0 not in dims:
            raise ValueError("bad empty() value")
        return np.zeros(dims, dtype=_TYPES[t])


# ---------------------------------------------------------------------------- binary reader
def _binary_value(buf: bytes, pos: int):
    if len(buf) < pos + 7:
        raise ValueError("truncated binary value")
    version, rank = buf[pos + 1], buf[pos + 2]
    if version != BINARY_VERSION:
        raise ValueError(f"binary format version {version} (expected {BINARY_VERSION})")
    name = bytes(buf[pos + 3:pos + 7])
    if name not in _BIN_TYPE:
        raise ValueError(f"unknown binary element type {name!r}")
    dt = _TYPES[_BIN_TYPE[name]]
    pos += 7
    shape = struct.unpack_from(f"<{rank}q", buf, pos) if rank else ()
    pos += 8 * rank
    if any(d < 0 for d in shape):
        raise ValueError("negative dimension")
    n = int(np.prod(shape, dtype=np.int64)) if rank else 1
    nbytes = n * dt.itemsize
    if len(buf) < pos + nbytes:
        raise ValueError("truncated binary value")
    a = np.frombuffer(buf, dtype=dt, count=n, offset=pos).copy()
    pos += nbytes
    if rank == 0:
        return a[0], pos
    return a.reshape(shape), pos


# ---------------------------------------------------------------------------- public API
def loads(data) -> list:
    """Every value in ``data`` (bytes or str; text and binary values may alternate)."""
    if isinstance(data, str):
        data = data.encode("utf-8")
    buf = bytes(data)
    out, pos, n = [], 0, len(buf)
    while pos < n:
        c = buf[pos:pos + 1]
        if c.isspace():
            pos += 1
            continue
        if c == b"b" and pos + 1 < n and buf[pos + 1] == BINARY_VERSION:
            v, pos = _binary_value(buf, pos)
            out.append(v)
            continue
        # a run of text up to the next binary value (text holds no byte 0x02)
        j = buf.find(b"b\x02", pos)
        if j < 0:
            j = n
        p = _Parser(buf[pos:j].decode("utf-8"))
        while not p.done():
            out.append(p.value())
        pos = j
    return out


def load(path: str) -> list:
    with open(path, "rb") as f:
        return loads(f.read())


def _text_scalar(v, t: str) -> str:
    if t == "bool":
        return "true" if bool(v) else "false"
    if t[0] == "f":
        x = float(v)
        if x != x:
            return f"{t}.nan"
        if x in (float("inf"), float("-inf")):
            return ("-" if x < 0 else "") + f"{t}.inf"
        return repr(x) + t
    return f"{int(v)}{t}"


def _text(v) -> str:
    a = np.asarray(v)
    t = type_name(a.dtype)
    if a.ndim == 0:
        return _text_scalar(a[()], t)
    if a.size == 0:
        return "empty(" + "".join(f"[{d}]" for d in a.shape) + t + ")"
    if a.ndim == 1:
        return "[" + ",".join(_text_scalar(x, t) for x in a.tolist()) + "]"
    return "[" + ",".join(_text(s) for s in a) + "]"


def _binary(v) -> bytes:
    a = np.asarray(v)
    t = type_name(a.dtype)
    a = np.array(a, dtype=_TYPES[t], order="C")  # keeps rank 0 (ascontiguousarray would not)
    head = b"b" + bytes([BINARY_VERSION, a.ndim]) + _BIN_NAME[t]
    return head + struct.pack(f"<{a.ndim}q", *a.shape) + a.tobytes()


def dumps(values: Iterable, binary: bool = True):
    """Values in order, binary (bytes) or text (str, one value per line)."""
    values = list(values)
    if binary:
        return b"".join(_binary(v) for v in values)
    return "".join(_text(v) + "\n" for v in values)


def dump(values: Iterable, path: str, binary: bool = True) -> None:
    data = dumps(values, binary)
    with open(path, "wb") as f:
        f.write(data if binary else data.encode("utf-8"))


# ---------------------------------------------------------------------------- running datasets
ENTRIES = ("forwardprojection", "backprojection", "forwardprojection_sparse",
           "backprojection_sparse", "SIRT", "SIRT3D")


def entry_point(name: str):
    """An instance of the drop-in class ``name`` (synkrotomo_amd.futhark.<name>.<name>)."""
    if name not in ENTRIES:
        raise ValueError(f"unknown entry point {name!r}; one of {', '.join(ENTRIES)}")
    import importlib
    mod = importlib.import_module(f"synkrotomo_amd.futhark.{name}")
    return getattr(mod, name)()


def run(name: str, values: list) -> np.ndarray:
    """Call entry point ``name`` with a dataset's values; the result on the host."""
    return entry_point(name).main(*values).get()


def _cli(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m synkrotomo_amd.futhark_data")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run", help="run an entry point on a dataset, print its result")
    r.add_argument("entry", choices=ENTRIES)
    r.add_argument("dataset")
    r.add_argument("-o", "--output", help="write the result here instead of stdout")
    r.add_argument("--binary", action="store_true", help="binary value format")
    b = sub.add_parser("bench", help="mean wall time per call (microseconds)")
    b.add_argument("entry", choices=ENTRIES)
    b.add_argument("dataset")
    b.add_argument("--runs", type=int, default=10)
    args = ap.parse_args(argv)
    values = load(args.dataset)
    ep = entry_point(args.entry)
    if args.cmd == "run":
        out = ep.main(*values).get()
        if args.output:
            dump([out], args.output, binary=args.binary)
        elif args.binary:
            sys.stdout.buffer.write(dumps([out], True))
        else:
            sys.stdout.write(dumps([out], False))
        return 0
    ep.main(*values).get()  # warm-up: geometry, kernels, caches
    t0 = time.perf_counter()
    for _ in range(args.runs):
        ep.main(*values).get()
    us = (time.perf_counter() - t0) / args.runs * 1e6
    print(json.dumps({"entry": args.entry, "dataset": args.dataset, "runs": args.runs,
                      "mean_us": round(us, 1)}))
    return 0


if __name__ == "__main__":
    sys.exit(_cli())
