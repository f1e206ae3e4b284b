"""Writers for small SEG-Y / SU files used by the ingest tests.

Header fields are placed by their SEG-Y rev 0 byte positions (1-based, as
the standard lists them), independently of the readers' word mapping.
"""
import numpy as np

from oracle import segy as osegy

# (first byte, size) of the 101 trace-header columns, SEG-Y rev 0
HEADER_FIELDS = ([(1 + 4 * i, 4) for i in range(7)] +          # bytes 1-28
                 [(29 + 2 * i, 2) for i in range(4)] +         # 29-36
                 [(37 + 4 * i, 4) for i in range(8)] +         # 37-68
                 [(69, 2), (71, 2)] +                          # 69-72
                 [(73, 4), (77, 4), (81, 4), (85, 4)] +        # source X/Y, group X/Y
                 [(89 + 2 * i, 2) for i in range(76)])         # 89-240
assert len(HEADER_FIELDS) == 101
COL_SRCX, COL_SRCY, COL_RECX, COL_RECY = 21, 22, 23, 24       # 0-based columns
COL_NS, COL_DT = 38, 39                                       # bytes 115-116, 117-118


def ibm_encode(x):
    """float -> IBM System/360 single (round to nearest fraction)."""
    x = np.asarray(x, np.float64)
    words = np.zeros(x.shape, np.uint32)
    nz = x != 0
    a = np.abs(x[nz])
    e = np.floor(np.log2(a) / 4.0).astype(np.int64) + 1        # a = f * 16**e, f in [1/16, 1)
    frac = np.round(a / 16.0 ** e * 2.0 ** 24).astype(np.int64)
    over = frac >= 2 ** 24
    frac[over] //= 16
    e[over] += 1
    sign = (x[nz] < 0).astype(np.int64)
    words[nz] = ((sign << 31) | ((e + 64) << 24) | frac).astype(np.uint32)
    return words


def ebcdic_encode(text):
    """ASCII -> EBCDIC through the inverse of the reader's table."""
    tbl = osegy.ebcdic_table()
    inv = {}
    for code in range(255, -1, -1):
        inv[int(tbl[code])] = code
    return bytes(inv.get(ord(ch), 0x6F) for ch in text)


def header_bytes(cols, big_endian):
    h = bytearray(240)
    for c, (b, size) in enumerate(HEADER_FIELDS):
        v = int(cols[c]) & ((1 << (8 * size)) - 1)
        h[b - 1:b - 1 + size] = v.to_bytes(size, "big" if big_endian else "little")
    return bytes(h)


def sample_bytes(samples, fmt, big_endian):
    bo = ">" if big_endian else "<"
    if fmt == 1:
        return samples.astype(bo + "u4").tobytes()
    if fmt == 2:
        return samples.astype(bo + "i4").tobytes()
    if fmt == 3:
        return samples.astype(bo + "i2").tobytes()
    return samples.astype(bo + "f4").tobytes()


def write_segy(path, samples, cols, fmt, *, big_endian=True, text=None, reel=None, su=False,
               dt_us=4000):
    """samples: [n, ns] already in the storage type (IBM words for format 1);
    cols: [n, 101] header values; reel: 27 values (defaults set ns, dt, format)."""
    n, ns = samples.shape
    with open(path, "wb") as f:
        if not su:
            text = text if text is not None else "C 1 CLIENT PSKTM TEST".ljust(80) * 40
            f.write(ebcdic_encode(text[:3200].ljust(3200)))
            r = np.zeros(27, np.int64) if reel is None else np.asarray(reel, np.int64).copy()
            if reel is None:
                r[5], r[7], r[9] = dt_us, ns, fmt
            b = bytearray(400)
            for i in range(3):
                b[4 * i:4 * i + 4] = int(r[i]).to_bytes(4, "big")
            for i in range(24):
                b[12 + 2 * i:14 + 2 * i] = int(r[3 + i]).to_bytes(2, "big")
            f.write(bytes(b))
        for k in range(n):
            f.write(header_bytes(cols[k], big_endian))
            f.write(sample_bytes(samples[k], fmt, big_endian))


def random_headers(rng, n, ns, dt_us=4000):
    cols = rng.integers(-30000, 30000, size=(n, 101)).astype(np.int64)
    cols[:, COL_SRCX] = rng.integers(-2_000_000, 2_000_000, n)     # negative coordinates too
    cols[:, COL_SRCY] = rng.integers(-2_000_000, 2_000_000, n)
    cols[:, COL_RECX] = rng.integers(-2_000_000_000, 2_000_000_000, n)
    cols[:, COL_RECY] = rng.integers(0, 2_000_000, n)
    cols[:, COL_NS] = ns
    cols[:, COL_DT] = dt_us
    return cols


def random_samples(rng, n, ns, fmt):
    x = rng.standard_normal((n, ns)) * 1000.0
    if fmt == 1:
        return ibm_encode(x)
    if fmt == 2:
        return rng.integers(-2**31, 2**31 - 1, size=(n, ns), dtype=np.int64).astype(np.int32)
    if fmt == 3:
        return rng.integers(-2**15, 2**15 - 1, size=(n, ns)).astype(np.int16)
    return x.astype(np.float32)
