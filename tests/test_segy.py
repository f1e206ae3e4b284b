"""SEG-Y / SU ingest: host readers vs the restatement of the reference's MATLAB
readers (oracle/segy.py), which is pinned against the writer's known values;
device decode and the file -> device loader vs the same restatement."""
import os

import numpy as np
import pytest

import oracle
from oracle import segy as osegy
import psktm
from segy_files import (COL_NS, COL_RECX, COL_SRCX, ebcdic_encode, random_headers,
                        random_samples, write_segy)


def make(tmp_path, fmt=1, n=37, ns=301, big_endian=True, seed=0, name="t.segy", **kw):
    rng = np.random.default_rng(seed)
    cols = random_headers(rng, n, ns)
    smp = random_samples(rng, n, ns, fmt)
    path = str(tmp_path / name)
    write_segy(path, smp, cols, fmt, big_endian=big_endian, **kw)
    return path, cols, smp


# ---------------------------- oracle pins --------------------------------

def test_oracle_ebcdic_known_codes():
    tbl = osegy.ebcdic_table()
    # EBCDIC (code page 037) letters, digits, space and punctuation
    for code, ch in [(0xC1, "A"), (0xC9, "I"), (0xD1, "J"), (0xE2, "S"), (0xE9, "Z"),
                     (0x81, "a"), (0xA9, "z"), (0xF0, "0"), (0xF9, "9"), (0x40, " "),
                     (0x60, "-"), (0x61, "/"), (0x7A, ":"), (0x7B, "#"), (0x7C, "@")]:
        assert chr(tbl[code]) == ch, hex(code)
    assert chr(tbl[0x4B]) == "?"          # '.' is not in the script's table


def test_oracle_reads_writer_values(tmp_path):
    path, cols, smp = make(tmp_path, fmt=2, n=9, ns=17)
    r = osegy.reel_header(path)
    assert r[7] == 17 and r[9] == 2 and r[5] == 4000
    assert osegy.trace_count(path) == 9
    h = osegy.trace_headers(path, 2, 5, saturate=False)
    assert np.array_equal(h, cols[2:7])
    hs = osegy.trace_headers(path, 2, 5, saturate=True)
    assert np.array_equal(hs, np.where(cols[2:7] < 0,
                                       np.where(np.isin(np.arange(101), [0, 1, 2, 3, 4, 5, 6,
                                                11, 12, 13, 14, 15, 16, 17, 18, 21, 22, 23, 24]),
                                                2**31 - 1, 2**15 - 1), cols[2:7]))
    assert np.array_equal(osegy.trace_samples(path, 3, 4), smp[3:7])
    assert osegy.text_header(path)[0].startswith("C 1 CLIENT PSKTM TEST")


def test_oracle_little_endian_words(tmp_path):
    path, cols, smp = make(tmp_path, fmt=3, n=4, ns=9, big_endian=False)
    assert np.array_equal(osegy.trace_headers(path, 0, 4, big_endian=False, saturate=False), cols)
    assert np.array_equal(osegy.trace_samples(path, 0, 4, big_endian=False), smp)


# ---------------------------- host readers -------------------------------

def test_text_and_reel_header(tmp_path):
    text = ("C 1 CLIENT: SURVEY-7 / LINE 42 #3 @ ACME  abc xyz ~{}" + "." * 10).ljust(80) * 40
    path, _, _ = make(tmp_path, text=text)
    lines = psktm.segy_text_header(path)
    assert lines == osegy.text_header(path)
    assert lines[0].startswith("C 1 CLIENT: SURVEY-7 / LINE 42 #3 @ ACME  abc xyz ~{}??????????")
    reel = list(range(1000, 1003)) + [7, 0, 4000, 4000, 301, 301, 1] + list(range(17))
    path2, _, _ = make(tmp_path, reel=reel, name="r.segy")
    assert np.array_equal(psktm.segy_reel_header(path2), osegy.reel_header(path2))
    assert list(psktm.segy_reel_header(path2)) == reel


@pytest.mark.parametrize("fmt", [1, 2, 3, 5])
def test_open_layout(tmp_path, fmt):
    path, _, _ = make(tmp_path, fmt=fmt, n=23, ns=77)
    L = psktm.segy_open(path)
    bps = 2 if fmt == 3 else 4
    assert (L.data_offset, L.n_traces, L.n_samples, L.format, L.bytes_per_sample) == \
        (3600, 23, 77, fmt, bps)
    assert L.record_bytes == 240 + 77 * bps and L.ragged == 0 and L.sample_interval_us == 4000
    assert L.n_traces == osegy.trace_count(path)
    with open(path, "ab") as f:          # a partial record at the end
        f.write(b"\0" * 100)
    L = psktm.segy_open(path)
    assert L.n_traces == 23 and L.ragged == 1


@pytest.mark.parametrize("big_endian", [True, False])
def test_trace_headers_host(tmp_path, big_endian):
    path, cols, _ = make(tmp_path, fmt=1, n=40, ns=50, big_endian=big_endian)
    L = psktm.segy_open(path, big_endian=big_endian)
    for sat in (False, True):
        got = psktm.segy_trace_headers(L, path, 5, 30, saturate=sat)
        ref = osegy.trace_headers(path, 5, 30, big_endian=big_endian, saturate=sat)
        assert np.array_equal(got, ref), sat
    assert (got[:, COL_SRCX] <= 2**31 - 1).all()
    got = psktm.segy_trace_headers(L, path, 0, 40)
    assert np.array_equal(got[:, COL_RECX], cols[:, COL_RECX])   # negative values intact


def test_errors(tmp_path):
    with pytest.raises(psktm.PsktmError):
        psktm.segy_open(str(tmp_path / "missing.segy"))
    path, _, _ = make(tmp_path, fmt=4, n=3, ns=8)
    with pytest.raises(psktm.PsktmError, match="format 4"):
        psktm.segy_open(path)
    path, _, _ = make(tmp_path, fmt=1, n=3, ns=8, name="ok.segy")
    L = psktm.segy_open(path)
    with pytest.raises(psktm.PsktmError, match="outside"):
        psktm.segy_trace_headers(L, path, 2, 2)


# ---------------------------- device decode ------------------------------

def expected_samples(smp, fmt, ibm_reference):
    if fmt == 1:
        if ibm_reference:
            return oracle.ibm32_to_ieee_ref(smp.reshape(-1)).reshape(smp.shape)
        w = smp.astype(np.int64)
        e = ((w >> 24) & 0x7F) - 64
        v = np.ldexp((w & 0xFFFFFF).astype(np.float64), 4 * e - 24)
        return np.where(w & 0x80000000, -v, v).astype(np.float32)
    return smp.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,big_endian,ibm_reference",
                         [(1, True, True), (1, True, False), (1, False, True), (2, True, True),
                          (3, True, True), (3, False, True), (5, True, True), (5, False, True)])
def test_decode_device(cuda, tmp_path, fmt, big_endian, ibm_reference):
    import torch
    path, cols, smp = make(tmp_path, fmt=fmt, n=61, ns=333 if fmt != 3 else 334,
                           big_endian=big_endian)
    L = psktm.segy_open(path, big_endian=big_endian)
    raw = np.fromfile(path, dtype=np.uint8)[3600:]
    rec = torch.as_tensor(raw).to(cuda)
    tr, hd = psktm.segy_decode(L, rec, L.n_traces, ibm_reference=ibm_reference)
    ref = expected_samples(smp, fmt, ibm_reference)
    assert np.array_equal(tr.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(hd.cpu().numpy(), osegy.trace_headers(path, 0, L.n_traces,
                                                                big_endian=big_endian,
                                                                saturate=False))


@pytest.mark.gpu
@pytest.mark.parametrize("chunk", [None, 3 * 1454 + 5])
def test_load_file_to_device(cuda, tmp_path, monkeypatch, chunk):
    path, cols, smp = make(tmp_path, fmt=1, n=50, ns=303)
    if chunk:                                   # several staging chunks
        monkeypatch.setenv("PSKTM_SEGY_CHUNK_BYTES", str(chunk))
    L = psktm.segy_open(path)
    tr, hd = psktm.segy_load(L, path, 7, 40, device=cuda)
    ref = oracle.ibm32_to_ieee_ref(smp[7:47].reshape(-1)).reshape(40, 303)
    assert np.array_equal(tr.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(hd.cpu().numpy(), cols[7:47])


@pytest.mark.gpu
def test_su_file(cuda, tmp_path):
    path, cols, smp = make(tmp_path, fmt=5, n=12, ns=64, big_endian=False, su=True,
                           name="line.su")
    L = psktm.segy_open(path, big_endian=False)
    assert (L.data_offset, L.n_traces, L.n_samples, L.format) == (0, 12, 64, 5)
    tr = psktm.segy_load(L, path, headers=False, device=cuda)
    assert np.array_equal(tr.cpu().numpy(), smp)
