"""Reader and writer for Futhark's value formats.

The reference's datasets (``finpar/LocVolCalib-data/*.in`` / ``*.out``,
``misc/ocean-sim/data/*``) are sequences of Futhark values in either the text
format (``0.03f32``, ``[1f32, 2f32]``, ``--`` line comments; README.md:43-51)
or the binary format (``'b'``, version 2, rank, a 4-character type name,
rank x u64 dims, then little-endian data; pbbs/pbbs2fut.c:13-23).  This module
reads both into numpy scalars/arrays and writes the text format, so datasets
can be passed straight to the entry points and results compared with the
expected outputs.
"""
from __future__ import annotations

import re
from typing import Iterable, List, Union

import numpy as np

Value = Union[np.ndarray, np.generic]

_BIN_TYPES = {
    b"  i8": np.int8, b" i16": np.int16, b" i32": np.int32, b" i64": np.int64,
    b"  u8": np.uint8, b" u16": np.uint16, b" u32": np.uint32, b" u64": np.uint64,
    b" f16": np.float16, b" f32": np.float32, b" f64": np.float64,
    b"bool": np.bool_,
}
_SUFFIX_TYPES = {
    "i8": np.int8, "i16": np.int16, "i32": np.int32, "i64": np.int64,
    "u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64,
    "f16": np.float16, "f32": np.float32, "f64": np.float64,
}
_TOKEN = re.compile(
    r"\s*(?:(\[)|(\])|(,)|(true|false)|"
    r"([-+]?(?:\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?|inf|nan)"
    r"(?:(i8|i16|i32|i64|u8|u16|u32|u64|f16|f32|f64))?))")


def _strip_comments(text: str) -> str:
    return "\n".join(line.split("--", 1)[0] for line in text.splitlines())


def _parse_text(text: str) -> List[Value]:
    text = _strip_comments(text)
    pos, n = 0, len(text)
    values: List[Value] = []

    def skip_ws(p: int) -> int:
        while p < n and text[p].isspace():
            p += 1
        return p

    def parse_value(p: int):
        m = _TOKEN.match(text, p)
        if not m:
            raise ValueError(f"bad Futhark value near: {text[p:p + 30]!r}")
        if m.group(1):  # array
            p = m.end()
            elems = []
            p = skip_ws(p)
            if p < n and text[p] == "]":
                return np.array([], dtype=np.float64), p + 1
            while True:
                v, p = parse_value(p)
                elems.append(v)
                p = skip_ws(p)
                if p < n and text[p] == ",":
                    p += 1
                    continue
                if p < n and text[p] == "]":
                    p += 1
                    break
                raise ValueError(f"expected , or ] near: {text[p:p + 30]!r}")
            if isinstance(elems[0], np.ndarray):
                return np.stack(elems), p
            dtypes = {np.asarray(e).dtype for e in elems}
            dt = np.result_type(*dtypes)
            return np.array(elems, dtype=dt), p
        if m.group(4):
            return np.bool_(m.group(4) == "true"), m.end()
        if m.group(5):
            lit, suf = m.group(5), m.group(6)
            if suf:
                lit = lit[: -len(suf)]
                return _SUFFIX_TYPES[suf](float(lit) if suf[0] == "f" else int(lit)), m.end()
            if re.fullmatch(r"[-+]?\d+", lit):
                return np.int32(int(lit)), m.end()
            return np.float64(float(lit)), m.end()
        raise ValueError(f"bad Futhark value near: {text[p:p + 30]!r}")

    while True:
        pos = skip_ws(pos)
        if pos >= n:
            break
        v, pos = parse_value(pos)
        values.append(v)
    return values


def _parse_binary(data: bytes) -> List[Value]:
    values: List[Value] = []
    pos = 0
    while pos < len(data):
        if data[pos:pos + 1].isspace():
            pos += 1
            continue
        if data[pos:pos + 1] != b"b":
            raise ValueError("not a Futhark binary value")
        version, rank = data[pos + 1], data[pos + 2]
        if version != 2:
            raise ValueError(f"unsupported binary version {version}")
        tname = data[pos + 3:pos + 7]
        dt = np.dtype(_BIN_TYPES[tname]).newbyteorder("<")
        pos += 7
        dims = np.frombuffer(data, dtype="<u8", count=rank, offset=pos).astype(np.int64)
        pos += 8 * rank
        count = int(np.prod(dims)) if rank else 1
        arr = np.frombuffer(data, dtype=dt, count=count, offset=pos).copy()
        pos += count * dt.itemsize
        values.append(arr.reshape(tuple(dims)) if rank else arr[0])
    return values


def loads(data: Union[str, bytes]) -> List[Value]:
    """Parse every value in a Futhark text or binary dataset."""
    if isinstance(data, bytes):
        stripped = data.lstrip()
        if stripped[:1] == b"b" and len(stripped) > 2 and stripped[1] == 2:
            return _parse_binary(stripped)
        data = data.decode("utf-8")
    return _parse_text(data)


def load(path: str) -> List[Value]:
    with open(path, "rb") as f:
        return loads(f.read())


def _fmt_scalar(v) -> str:
    v = np.asarray(v)
    if v.dtype == np.bool_:
        return "true" if bool(v) else "false"
    name = v.dtype.name.replace("float", "f").replace("uint", "u").replace("int", "i")
    if v.dtype.kind == "f":
        return f"{float(v)!r}{name}"
    return f"{int(v)}{name}"


def dumps(values: Iterable[Value]) -> str:
    """Write values in the Futhark text format, one per line."""
    out = []
    for v in values:
        a = np.asarray(v)
        if a.ndim == 0:
            out.append(_fmt_scalar(a))
        else:
            def rec(x):
                if x.ndim == 1:
                    return "[" + ", ".join(_fmt_scalar(e) for e in x) + "]"
                return "[" + ", ".join(rec(e) for e in x) + "]"
            out.append(rec(a))
    return "\n".join(out) + "\n"


def dumps_binary(values: Iterable[Value]) -> bytes:
    """Write values in the Futhark binary format: per value 'b', version 2,
    the rank, the 4-character type name, the dimensions as little-endian
    u64 and the elements little-endian (the layout pbbs/pbbs2fut.c:13-23
    writes and _parse_binary reads)."""
    names = {np.dtype(t): k for k, t in _BIN_TYPES.items()}
    out = bytearray()
    for v in values:
        a = np.asarray(v)
        if a.dtype not in names:
            raise ValueError(f"no Futhark binary type for {a.dtype}")
        out += b"b" + bytes([2, a.ndim]) + names[a.dtype]
        out += np.asarray(a.shape, dtype="<u8").tobytes()
        out += np.ascontiguousarray(a, dtype=a.dtype.newbyteorder("<")).tobytes()
    return bytes(out)


def dump_binary(values: Iterable[Value], path: str) -> None:
    with open(path, "wb") as f:
        f.write(dumps_binary(values))
