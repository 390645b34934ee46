"""Minimal CBOR (RFC 8949) encoder/decoder for the recording format's `.bin` files.

Covers the data model the recordings use: unsigned/negative integers, byte strings, text
strings, arrays (lists and tuples), maps with string keys, booleans, null and IEEE floats
(encoded as 64-bit; 16/32/64-bit accepted on decode), definite lengths on encode and
definite or indefinite lengths on decode.  Semantic tags are skipped on decode (the tagged
item is returned).  Written from the RFC; no third-party CBOR library is needed.
"""

from __future__ import annotations

import struct

import numpy as np

__all__ = ["dumps", "loads"]


def _head(major: int, n: int) -> bytes:
    if n < 24:
        return bytes([(major << 5) | n])
    if n < 1 << 8:
        return bytes([(major << 5) | 24, n])
    if n < 1 << 16:
        return bytes([(major << 5) | 25]) + struct.pack(">H", n)
    if n < 1 << 32:
        return bytes([(major << 5) | 26]) + struct.pack(">I", n)
    if n < 1 << 64:
        return bytes([(major << 5) | 27]) + struct.pack(">Q", n)
    raise OverflowError("integer too large for CBOR")


def _enc(obj, out: list) -> None:
    if obj is None:
        out.append(b"\xf6")
    elif obj is True:
        out.append(b"\xf5")
    elif obj is False:
        out.append(b"\xf4")
    elif isinstance(obj, (int, np.integer)):
        v = int(obj)
        out.append(_head(0, v) if v >= 0 else _head(1, -1 - v))
    elif isinstance(obj, (float, np.floating)):
        out.append(b"\xfb" + struct.pack(">d", float(obj)))
    elif isinstance(obj, (bytes, bytearray, memoryview)):
        b = bytes(obj)
        out.append(_head(2, len(b)))
        out.append(b)
    elif isinstance(obj, str):
        b = obj.encode("utf-8")
        out.append(_head(3, len(b)))
        out.append(b)
    elif isinstance(obj, (list, tuple)):
        out.append(_head(4, len(obj)))
        for x in obj:
            _enc(x, out)
    elif isinstance(obj, dict):
        out.append(_head(5, len(obj)))
        for k, v in obj.items():
            _enc(k, out)
            _enc(v, out)
    else:
        raise TypeError(f"cannot CBOR-encode {type(obj).__name__}")


def dumps(obj) -> bytes:
    out: list = []
    _enc(obj, out)
    return b"".join(out)


_BREAK = object()


class _Reader:
    def __init__(self, data: bytes):
        self.d = memoryview(data)
        self.i = 0

    def take(self, n: int) -> bytes:
        if self.i + n > len(self.d):
            raise ValueError("truncated CBOR data")
        b = bytes(self.d[self.i:self.i + n])
        self.i += n
        return b

    def arg(self, info: int) -> int | None:
        if info < 24:
            return info
        if info == 24:
            return self.take(1)[0]
        if info == 25:
            return struct.unpack(">H", self.take(2))[0]
        if info == 26:
            return struct.unpack(">I", self.take(4))[0]
        if info == 27:
            return struct.unpack(">Q", self.take(8))[0]
        if info == 31:
            return None  # indefinite length
        raise ValueError(f"reserved CBOR additional info {info}")

    def item(self):
        ib = self.take(1)[0]
        major, info = ib >> 5, ib & 31
        if major == 7:
            if info == 20:
                return False
            if info == 21:
                return True
            if info in (22, 23):
                return None
            if info == 25:
                return float(np.frombuffer(self.take(2), dtype=">f2")[0])
            if info == 26:
                return struct.unpack(">f", self.take(4))[0]
            if info == 27:
                return struct.unpack(">d", self.take(8))[0]
            if info == 31:
                return _BREAK
            if info < 24:
                return info  # unassigned simple value
            if info == 24:
                return self.take(1)[0]
            raise ValueError(f"unsupported CBOR simple value {info}")
        n = self.arg(info)
        if major == 0:
            return n
        if major == 1:
            return -1 - n
        if major in (2, 3):
            if n is None:
                parts = []
                while True:
                    x = self.item()
                    if x is _BREAK:
                        break
                    parts.append(x)
                return b"".join(parts) if major == 2 else "".join(parts)
            b = self.take(n)
            return b if major == 2 else b.decode("utf-8")
        if major == 4:
            out = []
            if n is None:
                while True:
                    x = self.item()
                    if x is _BREAK:
                        break
                    out.append(x)
            else:
                for _ in range(n):
                    out.append(self.item())
            return out
        if major == 5:
            out = {}
            if n is None:
                while True:
                    k = self.item()
                    if k is _BREAK:
                        break
                    out[k] = self.item()
            else:
                for _ in range(n):
                    k = self.item()
                    out[k] = self.item()
            return out
        if major == 6:
            return self.item()  # tag: return the tagged item
        raise ValueError(f"bad CBOR major type {major}")


def loads(data: bytes):
    r = _Reader(data)
    obj = r.item()
    if r.i != len(r.d):
        raise ValueError("trailing bytes after the CBOR item")
    return obj
