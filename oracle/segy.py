"""Restatement of the reference's SEG-Y readers (test infrastructure only).

* :func:`text_header` -- scripts/ktmGetInputMeta.m:30-56 (opt 1): EBCDIC reel
  header through the script's own lookup table, 40 lines of 80 characters.
* :func:`reel_header` -- ktmGetInputMeta.m:58-76 (opt 2): 3 x 4-byte and 24 x
  2-byte values, composed from bytes in file order (big-endian, unsigned)
  whatever the endianness argument.
* :func:`trace_headers` -- ktmGetInputMeta.m:78-153 (opt 3): 101 columns from
  the 120 16-bit words of each 240-byte trace header (7 int32, 4 int16, 8
  int32, 2 int16, 4 int32 -- source X/Y, receiver X/Y at columns 22-25 --
  then 76 int16). MATLAB's int32()/int16() saturate, so with
  ``saturate=True`` negative values read as 2147483647 / 32767 as the script
  does; ``saturate=False`` gives two's complement. The script's
  fseek(fileID, offset, 3600) passes 3600 as the origin argument; the intended
  position, 3600 + (first-1) * record, is used here.
* :func:`trace_samples` -- scripts/ktmGetInputTrace.m:27-75: raw sample words
  ('uint32' for format 1 and unknown formats, 'int32' for 2, 'int16' for 3).
* :func:`trace_count` -- scripts/ktmGetTotalTraceNo.m:18-32.

Ranges here are 0-based [first, first + count); the scripts take 1-based
inclusive ranges (and ktmGetInputTrace returns nothing for a one-trace range).
"""
from __future__ import annotations

import os

import numpy as np


def ebcdic_table():
    """tbl from ktmGetInputMeta.m:39-50 (1-based index = byte value); byte 0,
    which the script cannot index, maps to '?' like other unlisted codes."""
    tbl = [63] * 257
    def put(lo, chars):
        for i, ch in enumerate(chars):
            tbl[lo + i] = ch if isinstance(ch, int) else ord(ch)
    put(129, "abcdefghi")
    put(145, "jklmnopqr")
    put(161, "~stuvwxyz")
    put(192, "{ABCDEFGHI")
    put(208, "}JKLMNOPQR")
    put(224, ["\\", 0, "S", "T", "U", "V", "W", "X", "Y", "Z"])
    put(240, "0123456789")
    put(96, "-/")
    tbl[64] = ord(" ")
    tbl[122] = ord(":")
    tbl[123] = ord("#")
    tbl[124] = ord("@")
    return np.array(tbl[:256], dtype=np.uint8)   # index = byte value


def text_header(path):
    raw = np.fromfile(path, dtype=np.uint8, count=3200)
    chars = ebcdic_table()[raw]
    return [bytes(chars[80 * i:80 * (i + 1)]).decode("latin-1") for i in range(40)]


def reel_header(path):
    b = np.fromfile(path, dtype=np.uint8, count=3600)[3200:3600].astype(np.int64)
    out = [int((b[4 * i] << 24) | (b[4 * i + 1] << 16) | (b[4 * i + 2] << 8) | b[4 * i + 3])
           for i in range(3)]
    out += [int((b[12 + 2 * i] << 8) | b[13 + 2 * i]) for i in range(24)]
    return np.array(out, dtype=np.int64)


def bytes_per_sample(fmt):
    """ktmGetInputTrace.m:43-63: 2 bytes for format 3, else 4."""
    return 2 if fmt == 3 else 4


def trace_count(path):
    r = reel_header(path)
    tl = bytes_per_sample(r[9]) * r[7]
    return (os.path.getsize(path) - 3600) / (240 + tl)


def _words(path, first, count, ns, fmt, big_endian):
    rec = 240 + bytes_per_sample(fmt) * ns
    raw = np.fromfile(path, dtype=np.uint8, count=3600 + (first + count) * rec)[3600 + first * rec:]
    return raw.reshape(count, rec)


def trace_headers(path, first, count, big_endian=True, saturate=True):
    r = reel_header(path)
    ns, fmt = int(r[7]), int(r[9])
    raw = _words(path, first, count, ns, fmt, big_endian)[:, :240]
    w = raw.reshape(count, 120, 2).astype(np.int64)
    w = (w[:, :, 0] << 8) | w[:, :, 1] if big_endian else w[:, :, 0] | (w[:, :, 1] << 8)

    def i32(cols):                      # pairs of 16-bit words, in column order
        a = w[:, cols[0]:cols[1]].reshape(count, -1, 2)
        v = (a[:, :, 0] << 16) + a[:, :, 1] if big_endian else a[:, :, 0] + (a[:, :, 1] << 16)
        return np.minimum(v, 2**31 - 1) if saturate else v - ((v >> 31) << 32)

    def i16(cols):
        v = w[:, cols[0]:cols[1]]
        return np.minimum(v, 2**15 - 1) if saturate else v - ((v >> 15) << 16)

    out = np.concatenate([i32((0, 14)), i16((14, 18)), i32((18, 34)), i16((34, 36)),
                          i32((36, 44)), i16((44, 120))], axis=1)
    assert out.shape[1] == 101
    return out.astype(np.int64)


def trace_samples(path, first, count, big_endian=True):
    r = reel_header(path)
    ns, fmt = int(r[7]), int(r[9])
    raw = _words(path, first, count, ns, fmt, big_endian)[:, 240:]
    bo = ">" if big_endian else "<"
    dt = {2: bo + "i4", 3: bo + "i2"}.get(fmt, bo + "u4")
    return np.ascontiguousarray(raw).view(dt).reshape(count, ns)
